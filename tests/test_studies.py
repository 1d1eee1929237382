"""The physics-study driver (masterarbeit_amd/studies.py): its
autocorrelation estimator (LGT_Code/Autocorrelation/t_int.c:31-69) on series
with known answers, and the C1 study on the GPU."""
import numpy as np
import pytest

from masterarbeit_amd.studies import tau_int


def ar1(rho, n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    x[0] = rng.standard_normal()
    s = np.sqrt(1 - rho * rho)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + s * rng.standard_normal()
    return x


def test_tau_int_white_noise():
    x = np.random.default_rng(1).standard_normal(20000)
    tau, err = tau_int(x)
    assert tau == pytest.approx(0.5, abs=0.05)
    assert err == pytest.approx(np.sqrt(2 * tau) * x.std() / np.sqrt(len(x) - 1), rel=1e-3)


def test_tau_int_ar1_closed_form():
    """AR(1) with coefficient rho: tau_int = (1 + rho) / (2 (1 - rho))."""
    for rho in (0.5, 0.8):
        taus = [tau_int(ar1(rho, 20000, s))[0] for s in range(4)]
        exact = (1 + rho) / (2 * (1 - rho))
        assert np.mean(taus) == pytest.approx(exact, rel=0.15), (rho, taus)


def test_tau_int_constant_series():
    assert tau_int(np.full(100, 0.7)) == (0.5, 0.0)


def test_tau_int_sums_until_first_negative():
    """Restated by hand: Gamma_k with the two partial means of t_int.c:31-41."""
    x = np.array([1.0, 2.0, 0.5, 1.5, 3.0, 0.0, 2.0, 1.0])
    n = len(x)

    def gamma(k):
        a, b = x[:n - k], x[k:]
        return np.sum((a - a.mean()) * (b - b.mean())) / (n - k)

    t = 0.5 * gamma(0)
    k = 1
    while k < n and gamma(k) >= 0:
        t += gamma(k)
        k += 1
    tau, err = tau_int(x)
    assert tau == pytest.approx(t / gamma(0), rel=1e-14)
    assert err == pytest.approx(np.sqrt(2 * tau) * np.sqrt(gamma(0) / (n - 1)), rel=1e-14)


@pytest.mark.gpu
def test_study_c1_on_gpu():
    import masterarbeit_amd.engine as E
    from masterarbeit_amd import studies
    E.load_library()
    r = studies.c1()
    assert r["plaquette_per_sweep"][0] < 1.0
    assert r["plaquette_mean_50_100"] == pytest.approx(0.7325, abs=0.002)
    w = r["wloops"]
    assert len(w) == 36 and 0 < w["6,6"] < w["1,1"] < 1


def test_study_quaternion_helpers_match_the_link_algebra():
    """The finite-difference study's own product and su(2) exponential agree
    with the test suite's (tests/fields.py) on random quaternions."""
    import fields as F
    from masterarbeit_amd.studies import qmul, su2_exp
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, b = rng.normal(size=4), rng.normal(size=4)
        np.testing.assert_allclose(qmul(a, b), F.qmul(a, b), rtol=1e-14, atol=1e-14)
        X = np.concatenate([[0.0], rng.normal(size=3)])
        u = su2_exp(X, 0.3)
        assert np.linalg.norm(u) == pytest.approx(1.0, abs=1e-15)
        # exp(eps X) = cos(eps |x|) + sin(eps |x|) x/|x|, and d/deps at 0 is X
        d = (su2_exp(X, 1e-6) - su2_exp(X, -1e-6)) / 2e-6
        np.testing.assert_allclose(d, X, rtol=1e-8, atol=1e-8)


def test_reference_24_constants_match_fixture():
    """The reference's 24^4 equilibrium values quoted in the study driver
    (W_REF24, W_REF24_APE100) are those derived from the fixture of the
    reference's own series, to the digits quoted."""
    from masterarbeit_amd.studies import W_REF24, W_REF24_APE100
    from test_gpu_therm24 import reference_values
    ref = reference_values()
    for (R, T), (m, e) in W_REF24.items():
        rm, re = ref[f"{R},{T}"]
        assert m == pytest.approx(rm, rel=1e-7) and e == pytest.approx(re, rel=0.05), (R, T, rm, re)
    rm, re = ref["ape100"]
    assert W_REF24_APE100[0] == pytest.approx(rm, rel=1e-4) and W_REF24_APE100[1] == pytest.approx(re, rel=0.05)
