/*
 * compat_lifecycle.c -- the field-lifecycle functions of a reference-style
 * host program (the role of LGT_Code/modules/admin/init.c initGaugeField /
 * releaseGaugeField and modules/io/config_IO.c readConfig / writeConfig),
 * kept in their own object file as in the reference so that linking with
 * -Wl,--wrap=<name> routes the program's calls through liblgt_compat.so.
 * Test infrastructure: configurations are raw [VOL][4][4] doubles.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lgt_compat.h"

extern long driver_vol;
static double *g_field = NULL;

/* COMPAT_SITE_ORDER=reversed lays the sites out backwards in the allocation,
 * so that the pointer table is not the reference's contiguous one */
void initGaugeField(int start)
{
    (void)start;
    const char *order = getenv("COMPAT_SITE_ORDER");
    const int rev = order && !strcmp(order, "reversed");
    g_field = malloc(sizeof(double) * 16 * (size_t)driver_vol);
    for (long n = 0; n < driver_vol; n++)
        for (int d = 0; d < 4; d++) {
            const long slot = rev ? driver_vol - 1 - n : n;
            pu[n][d] = (su2mat *)(g_field + 16 * slot + 4 * d);
            pu[n][d]->c0 = 1.0;
            pu[n][d]->c1 = pu[n][d]->c2 = pu[n][d]->c3 = 0.0;
        }
}

void releaseGaugeField(void)
{
    free(g_field);
    g_field = NULL;
}

void readConfig(char *name)
{
    FILE *f = fopen(name, "rb");
    if (!f) exit(3);
    for (long n = 0; n < driver_vol; n++)
        for (int d = 0; d < 4; d++)
            if (fread(pu[n][d], sizeof(double), 4, f) != 4) exit(3);
    fclose(f);
}

void writeConfig(char *name)
{
    FILE *f = fopen(name, "wb");
    if (!f) exit(4);
    for (long n = 0; n < driver_vol; n++)
        for (int d = 0; d < 4; d++) fwrite(pu[n][d], sizeof(double), 4, f);
    fclose(f);
}
