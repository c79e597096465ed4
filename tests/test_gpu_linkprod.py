"""The MFMA-vs-VALU link-product experiment (wd_link_product): both kernels
form U h for every site and system from the same F16 inputs, checked against
the product computed in float64 from the F16-rounded inputs.  Products of F16
numbers are exact in FP32, so the only differences are FP32 accumulation
roundings: tolerance 1e-6 of the largest entry per repetition.  Site counts
that leave idle lanes (VALU: sites past the last of a workgroup) and idle
wavefronts (MFMA: site pairs past the last), and system counts below, at and
above the 8 columns of an MFMA tile."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(nsites, nsys, seed):
    rng = np.random.default_rng(seed)

    def c16(*shape):
        z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        return (z.real.astype(np.float16).astype(np.float64)
                + 1j * z.imag.astype(np.float16).astype(np.float64))
    return c16(nsites, 3, 3), c16(nsites, nsys, 2, 3)


@pytest.mark.parametrize("nsys", [1, 5, 8, 12, 17])
@pytest.mark.parametrize("inner", [1, 6])
def test_link_product_matches_reference(w, nsys, inner):
    nsites = 1030
    U, h = _inputs(nsites, nsys, 7 + nsys)
    ref = inner * np.einsum("sab,srpb->srpa", U, h)
    with w.Context([4, 4, 4, 4]) as ctx:
        got = {impl: ctx.link_product(impl, U, h, inner)[0]
               for impl in (w.LINKPROD_VALU, w.LINKPROD_MFMA)}
    scale = np.abs(ref).max()
    for impl, out in got.items():
        err = np.abs(out - ref).max() / scale
        print(f"impl={impl} nsys={nsys} inner={inner}: max error {err:.2e}", flush=True)
        assert err < 1e-6 * inner


def test_link_product_timing_and_errors(w):
    U, h = _inputs(4096, 12, 3)
    with w.Context([4, 4, 4, 4]) as ctx:
        for impl in (w.LINKPROD_VALU, w.LINKPROD_MFMA):
            out, ms = ctx.link_product(impl, U, h, inner=16, iters=3)
            assert ms > 0
        for args in ((2, U, h), (w.LINKPROD_VALU, U[:3], h[:3])):
            with pytest.raises(w.WilsonError) as e:
                ctx.link_product(*args)
            assert e.value.code == w.WD_EARG
        with pytest.raises(w.WilsonError) as e:
            ctx.link_product(w.LINKPROD_MFMA, U, h, inner=0)
        assert e.value.code == w.WD_EARG
        U2, h2 = _inputs(8, 65, 4)
        with pytest.raises(w.WilsonError) as e:
            ctx.link_product(w.LINKPROD_MFMA, U2, h2)
        assert e.value.code == w.WD_EARG
