"""The allocation-free entry point (lmff_pair_compute_noalloc: device arrays in and out, every
working array in the library's fixed per-GPU scratch) against the allocating host-buffer
Pair::compute (lmff_pair_compute) on the same atoms: two calls in a row, the second smaller,
then the first again (the scratch keeps the previous call's contents)."""
import numpy as np
import pytest
import torch

from lmff import capi, structures as S

pytestmark = pytest.mark.gpu


def _run(ctx, s, eflag, vflag, f0):
    x = torch.tensor(s.x, dtype=torch.float64, device="cuda")
    t = torch.tensor(s.type, dtype=torch.int32, device="cuda")
    tag = torch.tensor(s.tag, dtype=torch.int64, device="cuda")
    mol = torch.tensor(s.mol, dtype=torch.int64, device="cuda")
    f = torch.tensor(f0, dtype=torch.float64, device="cuda")
    g = torch.full((9,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    print(f"noalloc call: {s.n} atoms, eflag {eflag}, vflag {vflag}", flush=True)
    ctx.pair_compute_noalloc(s.n, x.data_ptr(), t.data_ptr(), tag.data_ptr(), mol.data_ptr(), s.boxlo, s.boxhi,
                             s.periodic, eflag, vflag, f.data_ptr(), g.data_ptr())
    torch.cuda.synchronize()
    return f.cpu().numpy(), g.cpu().numpy()


@pytest.mark.parametrize("precision", [capi.LMFF_FP64, capi.LMFF_MIXED])
def test_noalloc_matches_pair_compute(precision):
    big = S.layered_stack(30, 20, ["G", "BN"], seed=31)     # 4,800 atoms
    small = S.layered_stack(15, 12, ["G", "BN"], seed=32)   # 1,440 atoms
    rng = np.random.default_rng(5)
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, big, precision=precision)
        for s in (big, small, big):
            ref = ctx.pair_compute(s, 1, 2)
            f0 = rng.normal(size=(s.n, 3))
            f, g = _run(ctx, s, 1, 2, f0)
            df = (f - f0) - ref["f"]
            rms = np.sqrt((ref["f"] ** 2).sum(1).mean())
            if precision == capi.LMFF_FP64:
                # FP64: the bar of every FP64 parity test (only the summation order differs)
                assert np.abs(df).max() <= 1e-10 * rms
                tol = 1e-10
            else:
                # mixed: the FP32 normals may take their neighbours in another order
                assert np.sqrt((df ** 2).sum(1).mean()) <= 1e-6
                tol = 1e-7
            assert g[0] == pytest.approx(ref["energy"], rel=tol)
            assert g[7] == pytest.approx(ref["e_vdw"], rel=tol)
            assert g[8] == pytest.approx(ref["e_rep"], rel=tol)
            vs = np.abs(ref["virial"]).max()
            assert np.abs(g[1:7] - ref["virial"]).max() <= tol * vs
        # forces only: the globals come back as zeros
        f, g = _run(ctx, small, 0, 0, np.zeros((small.n, 3)))
        ref = ctx.pair_compute(small, 0, 0)
        assert np.abs(f - ref["f"]).max() <= (1e-10 if precision == capi.LMFF_FP64 else 1e-5)
        assert (g == 0).all()


def test_noalloc_refusals():
    s = S.layered_stack(15, 12, ["G", "BN"], seed=33)
    with capi.Context(0) as ctx:
        capi.setup_for(ctx, s)
        with pytest.raises(capi.LmffError) as e:
            _run(ctx, s, 3, 2, np.zeros((s.n, 3)))    # per-atom energy
        assert e.value.code == 1
        with pytest.raises(capi.LmffError) as e:
            _run(ctx, s, 1, 6, np.zeros((s.n, 3)))    # per-atom virial
        assert e.value.code == 1
        # more atoms than the fixed scratch holds: refused before any kernel runs
        assert capi.load().lmff_noalloc_scratch_bytes() == 256 << 20
        huge = S.layered_stack(150, 100, ["G", "BN"], seed=34, temperature=None)   # 120,000 atoms
        with pytest.raises(capi.LmffError) as e:
            _run(ctx, huge, 1, 2, np.zeros((huge.n, 3)))
        assert e.value.code == 8
        # the context still works afterwards
        f, g = _run(ctx, s, 1, 2, np.zeros((s.n, 3)))
        assert g[0] == pytest.approx(ctx.pair_compute(s, 1, 2)["energy"], rel=1e-10)
