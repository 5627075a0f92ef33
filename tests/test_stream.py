"""Out-of-core streamer: halo planning and lazy fields (CPU), slab-stitched = whole field (GPU)."""
import numpy as np
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import gen
from stenpy_amd.stream import (GeneratorField, NpyLazyField, composite_radius, halo_ranges, plan_slabs,
                               stream_graph)


def ref_halo_rows(start, end, total, radius, boundary):
    """Row indices the reference's _read_halo_chunk concatenates (stenpy.py:924-951)."""
    if radius <= 0:
        return list(range(start, end)), 0, 0
    if boundary != "periodic":
        lo, hi = max(0, start - radius), min(total, end + radius)
        return list(range(lo, hi)), start - lo, hi - end
    rows = [(start + d) % total for d in range(-radius, end - start + radius)]
    return rows, radius, radius


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_halo_ranges_match_reference(bc):
    total = 23
    for r in (0, 1, 2, 3):
        for start in range(0, total, 5):
            end = min(start + 5, total)
            pieces, lp, rp = halo_ranges(start, end, total, r, bc)
            rows = [i for a, b in pieces for i in range(a, b)]
            assert (rows, lp, rp) == ref_halo_rows(start, end, total, r, bc)


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_slab_plans_rebuild_the_reference_halo_rows(bc):
    """Every padded slab, assembled from its moves (host reads, copies from the previous slab's
    device buffer, the saved first planes), holds exactly the rows the reference reads."""
    for n0 in (1, 2, 5, 7, 16, 23, 40):
        for slab in (1, 2, 3, 5, 7, 16, 40):
            for r in (0, 1, 2, 3):
                if bc == "periodic" and r > n0:
                    continue
                plans = plan_slabs(n0, slab, r, bc)
                for pl in plans:
                    rows = [None] * pl.planes
                    for dst, kind, a, b in pl.moves:
                        for i in range(b - a):
                            if kind == "prev":
                                u = plans[pl.k - 1].lo + a + i
                                rows[dst + i] = u % n0 if bc == "periodic" else u
                            else:                    # "host" and "wrap" name global planes
                                rows[dst + i] = a + i
                    want, lp, rp = ref_halo_rows(pl.s, pl.e, n0, r, bc)
                    assert rows == want and pl.lpad == lp, (n0, slab, r, bc, pl.k)
                    assert not [m for m in pl.moves if m[1] == "prev"] or pl.k > 0


def test_composite_radius():
    f = torch.zeros(4, 4, 4)
    for expr, r in [("f*f", 0), ("laplacian(f)", 1), ("laplacian(laplacian(f))", 2),
                    ("f**2 + tanh(f)*exp(f) + laplacian(f)", 1), ("eigenvalues(hessian(f)) ", 2),
                    ("divergence(curl(gradient(f)))", 3)]:
        g, *_ = sp.compile_expression(expr, field_map={"f": f})
        assert composite_radius(g) == r, expr


def test_generator_field_matches_host_formula():
    lf = GeneratorField((30, 20, 18), 7)
    full = gen.generate_host((30, 20, 18), 7)
    np.testing.assert_array_equal(lf[5:17], full[5:17])
    out = np.empty((4, 20, 18), np.float32)
    lf.read_into(26, 30, out)
    np.testing.assert_array_equal(out, full[26:30])


def test_generator_field_rows_span_several_column_segments():
    """Rows longer than the generator's 256-column segments and row counts that leave a
    partial 16-row block: every voxel still equals the numpy formula."""
    shape = (5, 23, 601)
    lf = GeneratorField(shape, 11)
    full = gen.generate_host(shape, 11)
    np.testing.assert_array_equal(lf[1:4], full[1:4])
    np.testing.assert_array_equal(lf[0:5], full)


def test_npy_lazy_field(tmp_path):
    a = gen.generate_host((12, 8, 9), 3)
    p = tmp_path / "f.npy"
    np.save(p, a)
    lf = NpyLazyField(str(p))
    out = np.empty((3, 8, 9), np.float32)
    lf.read_into(4, 7, out)
    np.testing.assert_array_equal(out, a[4:7])


STREAM_EXPR = "f**2 + tanh(f)*exp(f) + laplacian(f)"


@pytest.mark.gpu
@pytest.mark.parametrize("stage_mb", [1024.0, 0.03])       # whole slabs / 3-plane staging chunks
@pytest.mark.parametrize("slab", [7, 16, 40])
def test_streamed_equals_whole_field(slab, stage_mb, tmp_path):
    shape = (40, 36, 70)
    lf = GeneratorField(shape, 7)
    whole = torch.from_numpy(lf[0:40]).cuda()
    g, sink, _, fm, _ = sp.compile_expression(STREAM_EXPR, boundary="periodic", field_map={"f": whole})
    ref = sp.Runtime(device="cuda").run(g)[sink].cpu()
    probe = torch.zeros((1,) + shape[1:])
    g2, sink2, _, fm2, _ = sp.compile_expression(STREAM_EXPR, boundary="periodic", field_map={"f": probe})
    outp = str(tmp_path / "out.npy")
    st = stream_graph(g2, sink2, fm2["f"], lf, out_path=outp, slab_planes=slab, stage_mb=stage_mb)
    got = torch.from_numpy(np.load(outp))
    assert torch.equal(got, ref), (got - ref).abs().max()
    l2 = float(torch.linalg.vector_norm(ref.double()))
    assert abs(st["l2"] - l2) <= 1e-6 * l2
    assert st["n_chunks"] == -(-40 // slab) and st["halo"] == 1
    assert st["stage_planes"] == (3 if stage_mb < 1 else slab + 2)


@pytest.mark.gpu
def test_producer_error_propagates():
    class Broken(GeneratorField):
        def read_into(self, z0, z1, out):
            if z0 > 10:
                raise IOError("disk on fire")
            super().read_into(z0, z1, out)

    lf = Broken((40, 16, 16), 1)
    probe = torch.zeros((1, 16, 16))
    g, sink, _, fm, _ = sp.compile_expression(STREAM_EXPR, boundary="periodic", field_map={"f": probe})
    with pytest.raises(RuntimeError, match="disk on fire"):
        stream_graph(g, sink, fm["f"], lf, slab_planes=8)


@pytest.mark.parametrize("distinct", [1, 2, 3, 4, 5, 6, 7, 8])
def test_host_generator_merges_modes_sharing_an_x_wavenumber(distinct):
    """Modes with the same x wavenumber share one term per voxel in the slab generator; every
    number of distinct wavenumbers (1..8, one code path each) gives the direct formula's
    values (cos per voxel and mode, numpy), to one ulp at most and almost always exactly."""
    from stenpy_amd import _lib as L
    ks, amps, phis = gen.fourier_params(5)
    ks = ks.copy()
    ks[:, 2] = [1 + (m % distinct) for m in range(8)]
    shape, z0, nz = (9, 7, 300), 2, 5
    want = gen.generate_host(shape, 5, z0=z0, nz=nz, modes=(ks, amps, phis))
    got = np.empty((nz,) + shape[1:], np.float32)
    L.host_gen_fourier(got.ctypes.data, shape, z0, nz, ks, amps, phis, 5, 2)
    np.testing.assert_array_max_ulp(got, want, maxulp=1)
    assert np.count_nonzero(got != want) <= 2
