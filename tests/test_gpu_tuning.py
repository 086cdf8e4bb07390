"""Launch-shape variants of the contraction give the same W (to summation
order), and the autotuner picks one of its candidates."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_contraction_variants_agree(wl, precision):
    dims, bc = (12, 6, 6, 8), 0
    u = O.generate(dims, bc, 44, 0.35)
    p = wl.make_parms([0, 1, 3, 2], 5, 3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms([0, 1, 3, 2], 5, 3))
    pk = wl.TUNE_MIXED_CONTRACT_KERNEL if precision else wl.TUNE_FP64_CONTRACT_KERNEL
    outs = []
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        for kernel, wgs in ((wl.KERNEL_STREAM, (1, 3, 64)), (wl.KERNEL_CHAIN, (1, 7, 512)),
                            (wl.KERNEL_CHAIN1, (1, 5, 256)), (wl.KERNEL_DEPTH, (1, 8, 40)),
                            (wl.KERNEL_CHAIN2, (1, 6, 512)), (wl.KERNEL_HYBRID, (1, 8, 40)),
                            (wl.KERNEL_STORED, (1, 8, 40)), (wl.KERNEL_STORED9, (1, 8, 40))):
            for wg in wgs:
                lat.tune(pk, kernel)
                lat.tune(pk + 1, wg)
                outs.append(lat.measure(p, precision))
    tol = 1e-12 if precision == 0 else 1e-6
    for w in outs:
        assert rel_err(w, ref) < tol
        assert rel_err(w, outs[0]) < (1e-13 if precision == 0 else 1e-7)


# fused time-slice forms (4: every T; 6: the FP32 blocks T <= 8 of the mixed
# path, the one-lane chain form beyond; 7: as 6, with the blocks from T = 10
# fused over the stored transporters): ragged pair blocks (spatial slice not a
# multiple of 16), start groups that overrun N0, open and periodic boundaries,
# and on the mixed path T blocks on both sides of the FP32/FP64 split at T = 9
# and stored blocks of one and of seven T
@pytest.mark.parametrize("dims,bc,mwlt,mwlr", [((10, 5, 6, 7), 0, 5, 2), ((20, 5, 6, 7), 3, 11, 3),
                                               ((18, 4, 4, 6), 0, 10, 2), ((36, 4, 4, 6), 3, 16, 2)])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kernel", [4, 6, 7, 8])
@pytest.mark.parametrize("traces", [1, 2])
def test_fused_time_slices(wl, dims, bc, mwlt, mwlr, precision, kernel, traces):
    nss = [0, 2, 3]
    u = O.generate(dims, bc, 45, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    pk, pw, pd, ps = ((wl.TUNE_MIXED_CONTRACT_KERNEL, wl.TUNE_MIXED_CONTRACT_WG, wl.TUNE_MIXED_CONTRACT_DEPTH,
                       wl.TUNE_MIXED_CONTRACT_STARTS) if precision else
                      (wl.TUNE_FP64_CONTRACT_KERNEL, wl.TUNE_FP64_CONTRACT_WG, wl.TUNE_FP64_CONTRACT_DEPTH,
                       wl.TUNE_FP64_CONTRACT_STARTS))
    tol = 1e-12 if precision == 0 else 1e-6
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        base = lat.measure(p, precision)  # library default form
        assert rel_err(base, ref) < tol
        lat.tune(pk, kernel)
        lat.tune(wl.TUNE_FUSED_TRACES, traces)
        for depth, starts, wg in ((16, 4, 8), (3, 4, 5), (1, 8, 1), (8, 8, 16)):
            lat.tune(pd, depth)
            lat.tune(ps, starts)
            lat.tune(pw, wg)
            print(f"[fused] kernel={kernel} traces={traces} dims={dims} bc={bc} prec={precision} "
                  f"depth={depth} starts={starts} wg={wg}", flush=True)
            w = lat.measure(p, precision)
            assert rel_err(w, ref) < tol, (depth, starts, wg)
            assert rel_err(w, base) < (1e-13 if precision == 0 else 1e-7), (depth, starts, wg)


# traces of the fused launches on the matrix cores (4x4 outer products per
# pair) against those on the vector units, for every level count the fused
# forms take (levels i >= msl of a pair's four lanes are idle) and an odd
# slice (ragged last pair block)
@pytest.mark.parametrize("nss", [[0], [0, 2], [0, 1, 2], [1, 1, 2, 3]])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kernel", [4, 7, 8])
def test_fused_traces_matrix_cores(wl, nss, precision, kernel):
    dims, bc, mwlt, mwlr = (26, 5, 6, 7), 3, 12, 3
    u = O.generate(dims, bc, 47, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    pk = wl.TUNE_MIXED_CONTRACT_KERNEL if precision else wl.TUNE_FP64_CONTRACT_KERNEL
    tol = 1e-12 if precision == 0 else 1e-6
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.tune(pk, kernel)
        lat.tune(wl.TUNE_FUSED_TRACES, 1)
        vec = lat.measure(p, precision)
        lat.tune(wl.TUNE_FUSED_TRACES, 2)
        print(f"[traces] kernel={kernel} nss={nss} prec={precision}", flush=True)
        mat = lat.measure(p, precision)
        again = lat.measure(p, precision)
    assert rel_err(vec, ref) < tol
    assert rel_err(mat, ref) < tol
    assert rel_err(mat, vec) < (1e-13 if precision == 0 else 1e-7)
    assert np.array_equal(mat, again)


# mixed path, fused form for every T with the blocks from T = 9 in FP32, each
# restarted from an FP32 copy of the FP64 time transporter at its first T
@pytest.mark.parametrize("dims,bc,mwlt,mwlr", [((20, 5, 6, 7), 3, 11, 3), ((18, 4, 4, 6), 0, 10, 2),
                                               ((24, 4, 4, 4), 3, 12, 2)])
def test_fused_restart(wl, dims, bc, mwlt, mwlr):
    nss = [0, 2, 3]
    u = O.generate(dims, bc, 46, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.tune(wl.TUNE_MIXED_CONTRACT_KERNEL, 4)
        lat.tune(wl.TUNE_MIXED_RESTART, 1)
        for depth, starts, wg in ((16, 8, 8), (3, 4, 5), (8, 8, 16)):
            lat.tune(wl.TUNE_MIXED_CONTRACT_DEPTH, depth)
            lat.tune(wl.TUNE_MIXED_CONTRACT_STARTS, starts)
            lat.tune(wl.TUNE_MIXED_CONTRACT_WG, wg)
            print(f"[restart] dims={dims} bc={bc} depth={depth} starts={starts} wg={wg}", flush=True)
            w = lat.measure(p, wl.MIXED)
            assert rel_err(w, ref) < 1e-6, (depth, starts, wg)
        lat.tune(wl.TUNE_LEAN, 1)
        lat.tune(wl.TUNE_MIXED_CONTRACT_DEPTH, 0)
        print(f"[restart] dims={dims} bc={bc} lean", flush=True)
        assert rel_err(lat.measure(p, wl.MIXED), ref) < 1e-6


def test_restart_rejects_bad_value(wl):
    with wl.Lattice((4, 4, 4, 4), 3) as lat:
        with pytest.raises(wl.WloopError, match="mixed restart"):
            lat.tune(wl.TUNE_MIXED_RESTART, 2)


# two-lane chain form: msl 2 and 4 (msl 3 runs the one-lane form), a spatial
# slice that is not a multiple of the 128-site block, chains of several
# lengths (T against N0 = 10), open and periodic boundaries.  Each trace has
# the bits of the one-lane form's trace; only the FP64 summation order differs.
@pytest.mark.parametrize("nss", [[0, 2], [0, 1, 3], [0, 2, 2, 5]])
@pytest.mark.parametrize("bc", [0, 3])
@pytest.mark.parametrize("precision", [0, 1])
def test_two_lane_chain(wl, nss, bc, precision):
    dims, mwlt, mwlr = (10, 5, 6, 7), 6, 3
    u = O.generate(dims, bc, 47, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    pk = wl.TUNE_MIXED_CONTRACT_KERNEL if precision else wl.TUNE_FP64_CONTRACT_KERNEL
    tol = 1e-12 if precision == 0 else 1e-6
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.tune(pk, wl.KERNEL_CHAIN1)
        lat.tune(pk + 1, 7)
        one = lat.measure(p, precision)
        lat.tune(pk, wl.KERNEL_CHAIN2)
        for wg in (1, 3, 64):
            lat.tune(pk + 1, wg)
            print(f"[two-lane] nss={nss} bc={bc} prec={precision} wg={wg}", flush=True)
            w = lat.measure(p, precision)
            assert rel_err(w, ref) < tol, wg
            assert rel_err(w, one) < 1e-11, wg


def test_fused_rejects_bad_shapes(wl):
    with wl.Lattice((8, 4, 4, 4), 3) as lat:
        for param, value in ((wl.TUNE_FP64_CONTRACT_KERNEL, 9), (wl.TUNE_FP64_CONTRACT_DEPTH, 17),
                             (wl.TUNE_FUSED_TRACES, 3),
                             (wl.TUNE_MIXED_CONTRACT_STARTS, 3), (wl.TUNE_FP64_CONTRACT_STARTS, 2)):
            with pytest.raises(wl.WloopError):
                lat.tune(param, value)


def test_autotune_picks_a_candidate(wl, tmp_path):
    from wloop_amd import autotune
    dims = (8, 8, 8, 8)
    p = wl.make_parms([0, 4, 4], 4, 4)
    with wl.Lattice(dims, 3) as lat:
        lat.generate(1, 0.4)
        cache = str(tmp_path / "tune.json")
        cands = [{"kernel": wl.KERNEL_STREAM, "wg": 2}, {"kernel": wl.KERNEL_CHAIN, "wg": 64},
                 {"kernel": wl.KERNEL_DEPTH, "wg": 8, "depth": 4, "starts": 4}]
        ch = autotune.tune(lat, p, wl.FP64, cache=cache, candidates=cands)
        assert (ch["kernel"], ch["wg"]) in {(1, 2), (2, 64), (4, 8)}
        assert ch["ms_per_rt"] > 0
        assert lat.tune_get(wl.TUNE_FP64_CONTRACT_KERNEL) == ch["kernel"]
        again = autotune.tune(lat, p, wl.FP64, cache=cache)
        assert again["cached"] and again["kernel"] == ch["kernel"]


@pytest.mark.parametrize("precision", [0, 1])
def test_ape_forms_agree(wl, precision):
    """APE with one thread per link (1) or per site (3), each as one kernel or as
    staple sums then projection (2, 4): one kernel and its two-kernel split do
    the same arithmetic, so the same bits; every form matches the oracle.  Odd
    and even extents, open bc."""
    dims, bc = (8, 6, 5, 7), 0
    u = O.generate(dims, bc, 46, 0.3)
    nss = [0, 2, 3]
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, 4, 2))
    p = wl.make_parms(nss, 4, 2)
    outs = {}
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        for form in (1, 2, 3, 4, 5):
            lat.tune(wl.TUNE_APE_FORM, form)
            outs[form] = lat.measure(p, precision)
    assert np.array_equal(outs[1], outs[2])
    assert np.array_equal(outs[3], outs[4])
    assert np.array_equal(outs[1], outs[5])  # packed FP32 layout (FP64: form 1 itself)
    for form in (1, 3):
        assert rel_err(outs[form], ref) < (1e-12 if precision == 0 else 1e-6), form


# APE on the packed FP32 layout: an odd volume (the packed buffers are aligned
# inside the scratch slots), a first level that is smeared (packed from the
# field) and levels of one step
@pytest.mark.parametrize("nss", [[1, 2, 1], [0, 1, 3]])
def test_ape_packed_layout(wl, nss):
    dims, bc = (9, 5, 5, 7), 3
    u = O.generate(dims, bc, 48, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, 3, 2))
    p = wl.make_parms(nss, 3, 2)
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.tune(wl.TUNE_APE_FORM, 1)
        w1 = lat.measure(p, wl.MIXED)
        lat.tune(wl.TUNE_APE_FORM, 5)
        w5 = lat.measure(p, wl.MIXED)
    assert np.array_equal(w1, w5)
    assert rel_err(w5, ref) < 1e-6


# lean layout (least device memory): only storage moves, so it gives the bits of
# the fused form with the same T depth; T blocks starting past T = 1 in both
# precisions (depth 4, T up to 11), level 0 aliasing the field (nss[0] = 0) or
# smeared, odd and even step counts per level, and a batch of lean contexts
@pytest.mark.parametrize("nss", [[0, 3, 2], [1, 2, 3]])
@pytest.mark.parametrize("precision", [0, 1])
def test_lean_layout(wl, nss, precision):
    dims, bc, mwlt, mwlr = (22, 5, 6, 7), 0, 11, 3
    u = O.generate(dims, bc, 48, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    pk, pd = ((wl.TUNE_MIXED_CONTRACT_KERNEL, wl.TUNE_MIXED_CONTRACT_DEPTH) if precision else
              (wl.TUNE_FP64_CONTRACT_KERNEL, wl.TUNE_FP64_CONTRACT_DEPTH))
    with wl.Lattice(dims, bc) as a, wl.Lattice(dims, bc) as b:
        a.upload(u)
        b.upload(u)
        a.tune(pk, wl.KERNEL_DEPTH)
        a.tune(pd, 4)
        b.tune(wl.TUNE_LEAN, 1)
        b.tune(pd, 4)
        fused = a.measure(p, precision)
        lean = b.measure(p, precision)
        assert np.array_equal(fused, lean)
        assert rel_err(lean, ref) < (1e-12 if precision == 0 else 1e-6)
        batch = wl.Batch([b], [p], precision)
        assert np.array_equal(batch.run()[0], lean)
        batch.close()


# forms 4 and 8 form the next R's spatial transporters inside the first fused
# launch of each R: a single R (nothing to form), several R with T blocks
# shallower than the T range (only the first launch of an R forms them), and
# the lean layout (one generation read while the next is written)
@pytest.mark.parametrize("mwlr,depth,lean", [(1, 0, 0), (4, 3, 0), (3, 0, 1)])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kernel", [4, 8])
def test_fused_forms_next_transporters(wl, mwlr, depth, lean, precision, kernel):
    dims, bc, mwlt, nss = (20, 8, 8, 9), 0, 10, [0, 1, 3]
    u = O.generate(dims, bc, 48, 0.3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    pk, pd = ((wl.TUNE_MIXED_CONTRACT_KERNEL, wl.TUNE_MIXED_CONTRACT_DEPTH) if precision else
              (wl.TUNE_FP64_CONTRACT_KERNEL, wl.TUNE_FP64_CONTRACT_DEPTH))
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.tune(pk, kernel)
        lat.tune(pd, depth)
        if lean:
            lat.tune(wl.TUNE_LEAN, 1)
        print(f"[next S] kernel={kernel} prec={precision} mwlr={mwlr} depth={depth} lean={lean}", flush=True)
        w = lat.measure(p, precision)
    assert rel_err(w, ref) < (1e-12 if precision == 0 else 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,wg", [((20, 4, 4, 4), 0), ((20, 6, 6, 6), 8)])
@pytest.mark.parametrize("starts", [0, 4])
def test_fused_sweep_start_groups(wl, dims, wg, starts):
    """The FP32 T <= 8 launch sweeping the start groups (d = W = 8, the mixed
    default) against the oracle and against the same launch without the sweep
    (4 start slices per workgroup): 20x4^3 has 12 pair blocks for 512
    workgroups (most own none and write zero sums); 20x6^3 with 8 workgroups
    takes ~5 pair blocks per workgroup (the first writes the partial slots,
    the others add); both over 3 start groups, the last ragged (20 = 8 + 8 +
    4); open bc, T = 1..10 (T = 9 chain, T = 10 over the stored
    transporters), R = 1, 2"""
    nss, mwlt, mwlr = [0, 1, 3], 10, 2
    u = O.generate(dims, 0, 51, 0.3)
    ref, _ = O.ms10(dims, 0, u, (1.0, 1.0, 0.5), O.make_parms(nss, mwlt, mwlr))
    p = wl.make_parms(nss, mwlt, mwlr)
    with wl.Lattice(dims, 0) as lat:
        lat.upload(u)
        lat.tune(wl.TUNE_MIXED_CONTRACT_WG, wg)
        lat.tune(wl.TUNE_MIXED_CONTRACT_STARTS, starts)
        print(f"[sweep] dims={dims} wg={wg} starts={starts}", flush=True)
        w = lat.measure(p, wl.MIXED)
        again = lat.measure(p, wl.MIXED)
    assert rel_err(w, ref) < 1e-6
    assert np.array_equal(w, again)  # the partial slots are written afresh by every launch
    assert np.array_equal(w == 0, ref == 0)
