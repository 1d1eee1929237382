/*
 * lgt_compat_wrap.c -- interceptors of a reference program's field-lifecycle
 * functions for the lazy host/device coherence of liblgt_compat.so.  Built
 * into liblgt_compat_wrap.a and linked statically into the program together
 * with -Wl,--wrap=<name> for each function below: the linker then sends the
 * program's calls to __wrap_<name> and resolves __real_<name> to the
 * program's own definition (LGT_Code/modules/admin/init.c:831-869 initGaugeField,
 * :897-905 copyGaugeField, :922-925 releaseGaugeField;
 * LGT_Code/modules/io/config_IO.c:40-117 writeConfig, :348-535
 * readConfigForThermalisation, :537-621 readConfig) and of the host-side
 * readers of pu outside the engine's path (obs/2pt.c measure2ptFunctions).  The __real_ references
 * are weak: a function the program does not wrap is never reached here.
 *
 * pu is marked changed before a function that writes it runs: readConfig
 * ends with its plaquette test, a call of the engine, which must then see
 * the new links (its upload clears the mark, so nothing is sent twice).
 */
#include "../../include/lgt_compat.h"

extern void __real_initGaugeField(int) __attribute__((weak));
extern void __real_releaseGaugeField(void) __attribute__((weak));
extern void __real_readConfig(char *) __attribute__((weak));
extern void __real_readConfigForThermalisation(char *) __attribute__((weak));
extern void __real_writeConfig(char *) __attribute__((weak));
extern void __real_copyGaugeField(su2mat *u1[][4], su2mat *u2[][4]) __attribute__((weak));
extern void __real_measure2ptFunctions(int *source, int stype, int n2pt, int *g1, int *g2) __attribute__((weak));

void __wrap_initGaugeField(int start)
{
    lgt_compat_lazy();
    lgt_compat_host_modified();
    __real_initGaugeField(start);
}

void __wrap_releaseGaugeField(void)
{
    lgt_compat_lazy();
    lgt_compat_host_modified();
    __real_releaseGaugeField();
}

void __wrap_readConfig(char *name)
{
    lgt_compat_lazy();
    lgt_compat_host_modified();
    __real_readConfig(name);
}

void __wrap_readConfigForThermalisation(char *name)
{
    lgt_compat_lazy();
    lgt_compat_host_modified();
    __real_readConfigForThermalisation(name);
}

void __wrap_writeConfig(char *name)
{
    lgt_compat_lazy();
    lgt_compat_sync_host();
    __real_writeConfig(name);
}

void __wrap_copyGaugeField(su2mat *u1[][4], su2mat *u2[][4])
{
    lgt_compat_lazy();
    if (u1 == pu) lgt_compat_sync_host();
    __real_copyGaugeField(u1, u2);
    if (u2 == pu) lgt_compat_host_modified();
}

/* the fermionic 2-pt functions (obs/2pt.c, modules.h:49) read pu on the host */
void __wrap_measure2ptFunctions(int *source, int stype, int n2pt, int *g1, int *g2)
{
    lgt_compat_lazy();
    lgt_compat_sync_host();
    __real_measure2ptFunctions(source, stype, n2pt, g1, g2);
}
