"""Tile choices of the single-electron orbital GEMM give identical bits.

The autotuner (tune_gemm) picks, per context, spin and batch size, the
workgroup size (4 or 2 waves) and the orbital tile per wave (16, 32, 64 or
128 orbitals) by timing.  Every output element accumulates the same k-chunks
in the same order whatever the tile, so the choice must not change a result
bit.  QMCB_GEMM_TILE fixes the tile; each valid tile is run on the same
walkers: a single-electron proposal (staged evaluate) and a few device VMC
steps, compared bit for bit, in every GEMM operand type (FP64, FP32, BF16
with its orbital-major coefficients) and with BF16 derivative storage.
"""
import numpy as np
import pytest

from qmcbeaver_amd import Context, synth

pytestmark = pytest.mark.gpu


def run(s, W, precision, tile, monkeypatch):
    monkeypatch.setenv("QMCB_GEMM_TILE", tile)
    R = synth.initial_walkers(s, W, 21)
    Rp = R.copy()
    Rp[:, 3, :] += 0.07
    ids = np.arange(W)
    with Context(s.pack(), precision=precision, max_walkers=W, seed=5) as c:
        c.evaluate(ids, R)
        p = c.evaluate(ids, Rp, whichE=3)
        c.evaluate(ids, R)
        st = c.vmc_steps(W, 0, 2 * s.n_elec, 0.02)
        return p, c.positions(W), c.sample_series(), st.sum_accept


@pytest.mark.parametrize("config,n_elec,tiles", [
    (1, None, ["4,2", "2,2", "4,1", "2,1"]),        # 32 orbitals per spin
    (-1, 100, ["4,4", "2,4", "4,2", "4,8", "2,8"]),  # 50 per spin, padded to 64: no 8
    (2, None, ["4,4", "2,4", "4,2", "4,8", "2,8"]),  # 128 per spin
])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16", "bf16store"])
def test_gemm_tiles_bit_identical(config, n_elec, tiles, precision, monkeypatch):
    s = synth.make_system(config, seed=7, n_elec=n_elec) if config < 0 else synth.make_system(config)
    n = s.n_up
    W = 256
    print("case GEMM tiles", config, n_elec, precision, flush=True)
    out = {}
    for tile in tiles:
        mt = int(tile.split(",")[1])
        pad = 16 if n <= 16 else 32 if n <= 32 else 64 if n <= 64 else 128 * ((n + 127) // 128)
        if -(-n // (16 * mt)) * 16 * mt > pad:
            with pytest.raises(Exception):
                run(s, W, precision, tile, monkeypatch)
            continue
        out[tile] = run(s, W, precision, tile, monkeypatch)
    assert len(out) >= 2
    ref_tile = next(iter(out))
    a = out[ref_tile]
    for tile, b in out.items():
        for k in ("log_psi", "e_local", "grad_lnpsi", "lap_ratio"):
            np.testing.assert_array_equal(a[0].a[k], b[0].a[k], err_msg=f"{tile} {k}")
        np.testing.assert_array_equal(a[1], b[1], err_msg=tile)
        np.testing.assert_array_equal(a[2], b[2], err_msg=tile)
        assert a[3] == b[3]
