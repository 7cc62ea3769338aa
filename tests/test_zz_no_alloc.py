"""rt_film_tonemap_noalloc (the entry point that allocates nothing: its
partial maxima live in a file-scope device array) against rt_film_tonemap,
through the C ABI.  Two films in a row, the second smaller and dimmer, so a
call that read the first call's partial maxima from the shared array would
map the second film with the wrong maxima and differ; every tone mapping,
both formats, with and without the render's valid counts.  The bar is byte
equality (integer output)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MODES = range(6)      # rt_tone_mapping: equalize .. gammaClamp
FORMATS = (0, 1)      # BMP, PPM


def tonemap(rt, ctx, fn, sums, nvalid, w, h, samples, mode, fmt, out):
    f = getattr(rt.lib(), fn)
    nv = rt._ptr(nvalid) if nvalid is not None else None
    ctx.launch(lambda s: rt.check(f(ctx.h, rt._ptr(sums), nv, w, h, samples, mode, fmt, rt._ptr(out), out.numel(),
                                    s)))


def test_noalloc_tonemap_equals_tonemap_on_two_calls_in_a_row(rt):
    ctx = rt.Context(0)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(17)
    for (w, h, spp, scale) in ((1920, 1080, 64, 8.0), (37, 23, 4, 0.25)):
        sums_np = rng.random((w * h, 3)) * scale * spp * rng.random((w * h, 1)) ** 2
        sums = torch.from_numpy(sums_np.reshape(-1)).to(dev)
        nvalid = torch.from_numpy(rng.integers(max(1, spp // 2), spp + 1, w * h).astype(np.int32)).to(dev)
        for nv in (None, nvalid):
            for fmt in FORMATS:
                nbytes = rt.lib().rt_image_pixel_bytes(fmt, w, h)
                for mode in MODES:
                    print(f"[no-alloc] {w}x{h} mode {mode} format {fmt} nvalid {nv is not None}", flush=True)
                    want = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
                    got = torch.full((nbytes,), 0xAB, dtype=torch.uint8, device=dev)
                    tonemap(rt, ctx, "rt_film_tonemap_noalloc", sums, nv, w, h, spp, mode, fmt, got)
                    tonemap(rt, ctx, "rt_film_tonemap", sums, nv, w, h, spp, mode, fmt, want)
                    torch.cuda.synchronize()
                    assert torch.equal(got, want), (w, h, mode, fmt)
    ctx.close()


def test_noalloc_tonemap_checks_its_arguments(rt):
    """The same refusals as rt_film_tonemap (no kernel runs)."""
    ctx = rt.Context(0)
    L = rt.lib()
    sums = torch.zeros(12, dtype=torch.float64, device="cuda")
    out = torch.zeros(64, dtype=torch.uint8, device="cuda")
    assert L.rt_film_tonemap_noalloc(ctx.h, rt._ptr(sums), None, 2, 2, 0, 5, 0, rt._ptr(out), 64, None) == \
        rt.RT_ERR_INVALID  # samples = 0
    assert L.rt_film_tonemap_noalloc(ctx.h, rt._ptr(sums), None, 2, 2, 1, 9, 0, rt._ptr(out), 64, None) == \
        rt.RT_ERR_INVALID  # unknown tone mapping
    assert L.rt_film_tonemap_noalloc(ctx.h, rt._ptr(sums), None, 2, 2, 1, 5, 0, rt._ptr(out), 3, None) == \
        rt.RT_ERR_INVALID  # pixel buffer too small
    assert b"rt_film_tonemap_noalloc" in L.rt_last_error()
    ctx.close()
