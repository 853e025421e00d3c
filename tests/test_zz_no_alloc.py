"""The allocation-free element-stiffness entry point
(ufc_ph_c3d8_eval_ke_noalloc) through the C-ABI: bitwise equal to
ufc_ph_c3d8_eval_ke on two calls in a row, the second with a smaller batch
(the static scratch keeps the first call's contents), for the incompatible-
mode element (the variant whose condensation factor lives in the scratch) and
the standard element, FP64 and FP32; batches beyond the scratch and host
arrays are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def _elements(n, seed):
    rng = np.random.default_rng(seed)
    xe = UNIT[None] * rng.uniform(0.5, 2.0, (n, 1, 3)) + 0.15 * rng.uniform(-1, 1, (n, 8, 3))
    return xe, rng.uniform(1e9, 2e11, n)


@pytest.mark.parametrize("variant", [4, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_eval_ke_noalloc_matches_eval_ke(ctx, variant, precision):
    import torch
    dev = torch.device("cuda:0")
    tdt = torch.float64 if precision == 0 else torch.float32
    for n, seed in ((9000, 1), (1234, 2)):
        print(f"[gpu-test] eval_ke_noalloc variant {variant} precision {precision} n {n}",
              flush=True)
        xe, E = _elements(n, seed)
        ref = ctx.eval_ke(xe, E, 0.3, precision=precision, variant=variant)
        xd = torch.as_tensor(xe, device=dev).contiguous()
        ed = torch.as_tensor(E, device=dev).contiguous()
        kd = torch.full((n, 24, 24), float("nan"), dtype=tdt, device=dev)
        ctx.eval_ke_noalloc(xd, ed, 0.3, kd, precision=precision, variant=variant)
        got = np.transpose(kd.cpu().numpy(), (0, 2, 1))
        assert np.array_equal(got, ref)


def test_eval_ke_noalloc_limits(ctx):
    import torch
    from ufc_hip import _abi as A
    dev = torch.device("cuda:0")
    n = 9710  # one element past the 16 MiB scratch for the incompatible-mode factor
    xd = torch.zeros((n, 8, 3), dtype=torch.float64, device=dev)
    ed = torch.ones(n, dtype=torch.float64, device=dev)
    kd = torch.zeros((n, 24, 24), dtype=torch.float64, device=dev)
    with pytest.raises(A.UfcError) as e:
        ctx.eval_ke_noalloc(xd, ed, 0.3, kd, variant=A.VARIANT_INCOMPAT)
    assert e.value.code == A.STATUS_INVALID and "exceeds the scratch" in e.value.message
    # host arrays would need staging copies: rejected
    import ctypes as C
    xe = np.zeros((4, 8, 3))
    E = np.ones(4)
    ke = np.zeros((4, 24, 24))
    algo = A.KeAlgo(0, 1)
    args = A.KeArgs(4, 0, xe.ctypes.data_as(C.c_void_p), E.ctypes.data_as(C.c_void_p), 0.3,
                    ke.ctypes.data_as(C.c_void_p))
    with pytest.raises(A.UfcError) as e:
        ctx._call("ufc_ph_c3d8_eval_ke_noalloc", C.byref(algo), C.byref(args))
    assert e.value.code == A.STATUS_INVALID and "device memory" in e.value.message
