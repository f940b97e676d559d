"""Pin the oracle's deterministic FP32 exp/log and the simulator's RNG against mathematics."""
import numpy as np

import oracle_lib as O
import qcstorm as Q


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def test_qexpf_accuracy():
    x = np.concatenate([np.linspace(-87.0, 88.0, 2_000_001), np.linspace(-1, 1, 200_001)]).astype(np.float32)
    y = np.zeros_like(x)
    O.lib().orc_qexpf_v(O.p(x), O.p(y), x.size)
    exact = np.exp(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    print("qexpf max ulp", err.max())
    assert err.max() <= 2.0


def test_qexpf_specials():
    e = O.lib().orc_qexpf
    assert e(-100.0) == 0.0 and e(0.0) == 1.0 and np.isinf(e(100.0)) and np.isnan(e(float("nan")))


def test_qlogf_accuracy():
    x = np.concatenate([np.geomspace(1.2e-38, 3e38, 2_000_001), np.linspace(0.5, 2.0, 400_001),
                        np.geomspace(1e-45, 1e-38, 10_001)]).astype(np.float32)
    x = x[x > 0]
    y = np.zeros_like(x)
    O.lib().orc_qlogf_v(O.p(x), O.p(y), x.size)
    exact = np.log(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    print("qlogf max ulp", err.max())
    assert err.max() <= 2.0


def test_qlogf_specials():
    l = O.lib().orc_qlogf
    assert np.isneginf(l(0.0)) and np.isnan(l(-1.0)) and np.isposinf(l(float("inf"))) and l(1.0) == 0.0


def _tlog(x):
    y = np.zeros_like(x)
    O.lib().orc_tlogf_v(O.p(x), O.p(y), x.size)
    return y


def test_tlogf_accuracy():
    """The fit loss's table log against double-precision log: every float of [1/4, 4) (all 32
    intervals in the binades around 1, where the table and polynomial terms dominate), a
    geometric sweep of the whole normal range, and subnormals."""
    x = np.arange(0x3e800000, 0x40800000, dtype=np.uint32).view(np.float32)
    y = _tlog(x)
    exact = np.log(x.astype(np.float64))
    nz = exact != 0.0
    err = np.abs(y[nz].astype(np.float64) - exact[nz]) / _ulp(exact[nz])
    print("tlogf max ulp on [1/4, 4)", err.max())
    assert err.max() <= 1.3 and y[x == 1.0][0] == 0.0
    x = np.concatenate([np.geomspace(1.2e-38, 3e38, 2_000_001), np.geomspace(1e-45, 1e-38, 10_001)]).astype(np.float32)
    x = x[(x > 0) & (x != 1.0)]
    y = _tlog(x)
    exact = np.log(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    print("tlogf max ulp, full range", err.max())
    assert err.max() <= 1.3


def test_tlog_table_entries():
    """Each pair (i_j, t_j): i_j within 0.1% of the reciprocal of its interval's centre, t_j the
    float nearest -log(i_j) to within 1e-3 ulp, the pair (1, 0) on the interval holding 1, and
    |m i_j - 1| < 2^-6 + 4e-4 over the interval's ends."""
    tab = np.zeros(64, np.uint32)
    O.lib().orc_tlog_table(O.p(tab))
    inv, t = tab[0::2].view(np.float32).astype(np.float64), tab[1::2].view(np.float32).astype(np.float64)
    k = np.uint32(0x3f3504f3)
    for j in range(32):
        lo = np.array([k + (j << 18)], np.uint32).view(np.float32)[0].astype(np.float64)
        hi = np.array([k + ((j + 1) << 18) - 1], np.uint32).view(np.float32)[0].astype(np.float64)
        if lo <= 1.0 <= hi:
            assert inv[j] == 1.0 and t[j] == 0.0 and j == 18
            continue
        assert abs(inv[j] * (lo + hi) / 2 - 1.0) < 1e-3, j
        assert abs(t[j] + np.log(inv[j])) <= 1e-3 * np.spacing(np.float32(abs(t[j]))), j
        assert max(abs(lo * inv[j] - 1.0), abs(hi * inv[j] - 1.0)) < 2.0 ** -6 + 4e-4, j


def test_tlog_u_adds_the_data_term():
    """orc_tlog_u(x, c) = log x + c with c = 1 - log q (the loss's data term) for model values x
    near and far from q: within 1 ulp of log x plus 1 ulp of the result, and c = 0 is tlogf."""
    rng = np.random.default_rng(5)
    q = rng.uniform(1.0, 5000.0, 20000).astype(np.float32)
    x = (q * rng.uniform(0.25, 4.0, q.size)).astype(np.float32)
    c = (1.0 - np.log(q.astype(np.float64))).astype(np.float32)
    u = np.array([O.lib().orc_tlog_u(float(a), float(b)) for a, b in zip(x, c)], np.float32)
    exact = np.log(x.astype(np.float64)) + c.astype(np.float64)
    bound = _ulp(np.log(x.astype(np.float64))) + _ulp(exact)
    assert np.all(np.abs(u.astype(np.float64) - exact) <= bound)
    assert np.array_equal(np.array([O.lib().orc_tlog_u(float(a), 0.0) for a in x[:200]], np.float32), _tlog(x[:200]))


def test_tlogf_specials():
    l = O.lib().orc_tlogf
    assert np.isneginf(l(0.0)) and np.isnan(l(-1.0)) and np.isposinf(l(float("inf"))) and l(1.0) == 0.0
    assert np.isnan(l(float("nan"))) and np.isneginf(O.lib().orc_tlog_u(0.0, 3.0))


def _tlogb(x):
    y = np.zeros_like(x)
    O.lib().orc_tlogbf_v(O.p(x), O.p(y), x.size)
    return y


def test_tlogbf_accuracy():
    """The one-emitter losses' binade table log against double-precision log: every float of
    [1/4, 4), every 7th float of [2^-11, 2^22) (the table's binades and their edges), a
    geometric sweep of the whole range (the scaled path outside the table) and subnormals."""
    for lo, hi, step, bound in ((0x3E800000, 0x40800000, 1, 1.05), (0x3A000000, 0x4A800000, 7, 1.05)):
        x = np.arange(lo, hi, step, dtype=np.uint32).view(np.float32)
        y = _tlogb(x)
        exact = np.log(x.astype(np.float64))
        nz = exact != 0.0
        err = np.abs(y[nz].astype(np.float64) - exact[nz]) / _ulp(exact[nz])
        print("tlogbf max ulp on [%g, %g)" % (x[0], x[-1]), err.max())
        assert err.max() <= bound and y[x == 1.0].tolist() in ([], [0.0])
    x = np.concatenate([np.geomspace(1.2e-38, 3e38, 2_000_001), np.geomspace(1e-45, 1e-38, 10_001)]).astype(np.float32)
    x = x[(x > 0) & (x != 1.0)]
    y = _tlogb(x)
    exact = np.log(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    print("tlogbf max ulp, full range", err.max())
    assert err.max() <= 1.05


def test_tlogb_table_entries():
    """Each pair (i, t) of binade e, interval j (128 per binade): i within 0.1% of the reciprocal
    of the interval's centre, t the float nearest -log(i) to within 1e-3 ulp,
    |x i - 1| < 2^-8 + 4e-4 over the interval's ends, and the interval holding 1 (binade 0,
    interval 73) exactly (1, 0)."""
    tab = np.zeros(8192, np.uint32)
    O.lib().orc_tlogb_table(O.p(tab))
    inv, t = tab[0::2].view(np.float32).astype(np.float64), tab[1::2].view(np.float32).astype(np.float64)
    k = 0x3F368000
    for idx in range(4096):
        g = idx - 1280
        lo_b = (k + (g << 16)) & 0xFFFFFFFF
        lo = float(np.array([lo_b], np.uint32).view(np.float32)[0])
        hi = float(np.array([lo_b + (1 << 16) - 1], np.uint32).view(np.float32)[0])
        if lo <= 1.0 <= hi:
            assert inv[idx] == 1.0 and t[idx] == 0.0 and idx == 1280 + 73
            continue
        assert abs(inv[idx] * (lo + hi) / 2 - 1.0) < 1e-3, idx
        assert abs(t[idx] + np.log(inv[idx])) <= 1e-3 * np.spacing(np.float32(abs(t[idx]))), idx
        assert max(abs(lo * inv[idx] - 1.0), abs(hi * inv[idx] - 1.0)) < 2.0 ** -8 + 4e-4, idx


def test_tlogb_u_adds_the_data_term():
    """orc_tlogb_u(x, c) = log x + c with c = 1 - log q for model values near and far from q,
    inside and outside the table's binades: within 1 ulp of log x plus 1 ulp of the result;
    c = 0 is tlogbf."""
    rng = np.random.default_rng(6)
    q = rng.uniform(1.0, 5000.0, 20000).astype(np.float32)
    x = (q * np.exp(rng.uniform(-12.0, 12.0, q.size))).astype(np.float32)
    c = (1.0 - np.log(q.astype(np.float64))).astype(np.float32)
    u = np.array([O.lib().orc_tlogb_u(float(a), float(b)) for a, b in zip(x, c)], np.float32)
    exact = np.log(x.astype(np.float64)) + c.astype(np.float64)
    bound = _ulp(np.log(x.astype(np.float64))) + _ulp(exact)
    assert np.all(np.abs(u.astype(np.float64) - exact) <= bound)
    assert np.array_equal(np.array([O.lib().orc_tlogb_u(float(a), 0.0) for a in x[:300]], np.float32), _tlogb(x[:300]))


def test_tlogbf_specials():
    l = O.lib().orc_tlogbf
    assert np.isneginf(l(0.0)) and np.isnan(l(-1.0)) and np.isposinf(l(float("inf"))) and l(1.0) == 0.0
    assert np.isnan(l(float("nan"))) and np.isneginf(O.lib().orc_tlogb_u(0.0, 3.0))


def _bits(v):
    return int(np.array([v], np.float32).view(np.uint32)[0])


def test_crlb_division_form_equals_ieee_quotient():
    """The CRLB's narrow-range division x * 50 + fma residual equals x / 0.02f for every float
    with |x| in [1e-30, 1e30] (the form is odd in x, so the positive half decides), and differs
    outside that range, where the kernel keeps the division."""
    lib = O.lib()
    assert lib.orc_div_dh_scan(_bits(1e-30), _bits(1e30) + 1) == 0
    assert lib.orc_div_dh_fast(0.0) == 0.0 and lib.orc_div_dh_fast(-0.0) == 0.0
    # the range is not vacuous: the form fails somewhere below it and above it
    assert lib.orc_div_dh_scan(1, _bits(1e-30)) > 0
    assert lib.orc_div_dh_scan(_bits(1e30) + 1, _bits(float("inf")) + 1) > 0
    # spot values against numpy's IEEE float32 quotient, negative side included
    rng = np.random.default_rng(5)
    x = (rng.uniform(-30, 30, 2000).astype(np.float64))
    x = (np.sign(rng.uniform(-1, 1, 2000)) * 10.0 ** x).astype(np.float32)
    want = x / np.float32(0.02)
    got = np.array([lib.orc_div_dh_fast(float(v)) for v in x], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10
    import ctypes as C
    # exercised through the simulator's public generator: first draw of stream (a, b, tag) = counter
    # (0, a, b, tag) -- checked against an independent numpy implementation of the round function
    def philox(ctr, key):
        c = [np.uint64(v) for v in ctr]
        k0, k1 = np.uint64(key[0]), np.uint64(key[1])
        M = np.uint64(0xFFFFFFFF)
        for _ in range(10):
            p0 = np.uint64(0xD2511F53) * c[0]
            p1 = np.uint64(0xCD9E8D57) * c[2]
            c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & M, p1 & M, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & M, p0 & M]
            k0 = (k0 + np.uint64(0x9E3779B9)) & M
            k1 = (k1 + np.uint64(0xBB67AE85)) & M
        return [int(v) for v in c]
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_simulator_poisson_statistics():
    # flat ROI without emitters: pixel electrons are Poisson(bg); mean and variance match
    sp = Q.default_sim(width=7, height=7, photons_min=1e-3, photons_max=1e-3, bg_min=37.0, bg_max=37.0,
                       kadc=1.0, offset=0.0)
    rois, _ = Q.sim_rois_host(sp, 5, 20000)
    pix = rois[:, :49].astype(np.float64).ravel()
    assert abs(pix.mean() - 37.0) < 0.05 and abs(pix.var() / 37.0 - 1) < 0.02
    sp2 = Q.default_sim(width=7, height=7, photons_min=1e-3, photons_max=1e-3, bg_min=4.0, bg_max=4.0,
                        kadc=1.0, offset=0.0)
    pix2 = Q.sim_rois_host(sp2, 6, 20000)[0][:, :49].astype(np.float64).ravel()
    assert abs(pix2.mean() - 4.0) < 0.02 and abs(pix2.var() / 4.0 - 1) < 0.02
