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
