"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Bars: bit-exact cell keys and sort permutation; FP64
fields within 1e-12 of the field maximum; FP32-compensated and FP32 within
1e-6 of the field maximum (tolerances written per test)."""
import ctypes

import numpy as np
import pytest
import torch

import bruteforce as bf
from ppm_legacy_amd import workloads as wl

pytestmark = pytest.mark.gpu

MP4, BSP2 = 2, 1
F64, F32C, F32 = 0, 1, 2
# F32 (the reference's single specifics) is checked against the float oracle,
# which sums in plain float and, like the GPU form, computes cell offsets in
# float (absolute offset error up to N * 2^-24); 1e-5 of the field maximum
# covers both.  F32C is checked against the FP64 oracle at the 1e-6 bar.
TOL = {F64: 1e-12, F32C: 1e-6, F32: 1e-5}


@pytest.fixture(scope="module")
def core():
    from ppm_legacy_amd import core as c
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return c


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def _pos_dtype(prec):
    return torch.float32 if prec == F32 else torch.float64


def _val_dtype(prec):
    return torch.float64 if prec == F64 else torch.float32


def _np_val(prec):
    return np.float64 if prec == F64 else np.float32


def _case(dim, N, n, seed, lda, lo=0.0, L=1.0):
    rng = np.random.default_rng(seed)
    xp = lo + L * rng.uniform(0, 1, size=(n, dim))
    up = rng.normal(size=(n, lda))
    return xp, up


# ------------------------------------------------------------ config 1 --
def test_p2m_config1_64cubed_fp64(core, oracle):
    """Plain 3D M'4 P2M, FP64, 64^3 periodic, 262,144 jittered particles (seed 1)."""
    N, g = 64, [2, 2, 2]
    xp = wl.jittered_lattice(N, seed=1)
    up = wl.strengths(1, len(xp), 1)
    t = oracle.topo(3, [N + 1] * 3)
    info, ref = oracle.p2m(t, MP4, xp, up, g)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    field = torch.zeros(mesh.padded_shape(1, g), dtype=torch.float64, device="cuda")
    plan = core.p2m(mesh, _dev(xp, torch.float64), _dev(up, torch.float64), field, [-2, -2, -2], MP4, F64)
    got = field.cpu().numpy()
    # every stored node: real nodes folded, ghosts raw, as the reference leaves them
    assert _rel(got, ref) < 1e-12
    assert _rel(oracle.real_nodes(got, g, 3), oracle.real_nodes(ref, g, 3)) < 1e-12
    uniq = oracle.unique_nodes(got, g, 3)
    assert abs(uniq.sum() - up.sum()) < 1e-9 * np.abs(up).sum()
    # cell keys and permutation: bit-exact against the reference's index arithmetic
    key, perm, start = plan.sort_view()
    x0 = (xp - 0.0) * 64.0
    cells = x0.astype(np.int64)                 # INT() truncation, ppm_interp_p2m.f:804-822
    ref_key = cells[:, 0] + N * (cells[:, 1] + N * cells[:, 2])
    assert np.array_equal(key.cpu().numpy(), ref_key)
    assert np.array_equal(perm.cpu().numpy(), np.argsort(ref_key, kind="stable"))
    assert np.array_equal(start.cpu().numpy(), np.concatenate([[0], np.cumsum(np.bincount(ref_key, minlength=N ** 3))]))


def test_binning_sorted_and_unsorted_input(core, oracle):
    """The binning has two paths: input already in cell order (cell heads
    written directly, identity permutation) and any other order (atomic slots,
    scan, scatter, stabilisation of the listed cells); a plan predicts the path
    from its last input.  Every (predicted, actual) pair -- one plan fed sorted, sorted,
    shuffled, shuffled, sorted input -- must give numpy's stable argsort, the
    exact cell heads and the oracle's field, on clustered particles with empty
    and crowded cells."""
    N, g = 24, [2, 2, 2]
    rng = np.random.default_rng(77)
    base = np.concatenate([rng.normal(0.3, 0.02, size=(3000, 3)) % 1.0, rng.uniform(0, 1, size=(1000, 3))])
    t = oracle.topo(3, [N + 1] * 3)
    mesh = core.Mesh(3, N + 1)
    plan = core.Plan()
    for order in ["cell_sorted", "cell_sorted", "shuffled", "shuffled", "cell_sorted"]:
        cells = (base * N).astype(np.int64)
        key = cells[:, 0] + N * (cells[:, 1] + N * cells[:, 2])
        o = np.argsort(key, kind="stable") if order == "cell_sorted" else rng.permutation(len(base))
        xp, key = base[o], key[o]
        up = rng.normal(size=(len(xp), 3))
        field = torch.zeros(mesh.padded_shape(3, g), dtype=torch.float64, device="cuda")
        core.p2m(mesh, _dev(xp, torch.float64), _dev(up, torch.float64), field, [-2] * 3, MP4, F64, plan)
        k, perm, start = plan.sort_view()
        assert np.array_equal(k.cpu().numpy(), key), order
        assert np.array_equal(perm.cpu().numpy(), np.argsort(key, kind="stable")), order
        assert np.array_equal(start.cpu().numpy(),
                              np.concatenate([[0], np.cumsum(np.bincount(key, minlength=N ** 3))])), order
        _, ref = oracle.p2m(t, MP4, xp, up, g)
        assert _rel(field.cpu().numpy(), ref) < 1e-12, order


@pytest.mark.parametrize("case", ["lattice", "long_gaps_overflow", "head_tail_chunks", "one_cell"])
def test_binning_direct_heads(core, oracle, case):
    """Input in cell order: the binning writes the cell heads straight from
    the keys (runs of up to 32 empty cells by the particle's lane, longer runs
    queued in pieces of 65536 cells, a full queue falls back to counting) and
    leaves the permutation implicit.  Cases: a jittered lattice (one particle
    per cell); 6000 particles 40 cells apart (more long gaps than the queue
    holds); particles only in z planes 30..33 of 64 (head and tail runs of
    over 65536 cells); all particles in one cell.  The plan is fed the same
    input twice (first call predicts unsorted input, second sorted): heads,
    identity permutation and keys must equal numpy's, the tiled P2M must
    match the oracle and repeat bitwise, and the tiled M2P must match the
    oracle in the caller's order."""
    N = 16 if case == "lattice" else 64
    rng = np.random.default_rng({"lattice": 1, "long_gaps_overflow": 2, "head_tail_chunks": 3, "one_cell": 4}[case])
    if case == "lattice":
        cells = np.arange(N ** 3)
    elif case == "long_gaps_overflow":
        cells = np.arange(0, N ** 3, 40)[:6000]
    elif case == "head_tail_chunks":
        cells = np.sort(rng.integers(30 * N * N, 34 * N * N, size=3000))
    else:
        cells = np.full(500, 100000)
    ijk = np.stack([cells % N, (cells // N) % N, cells // (N * N)], 1)
    xp = (ijk + rng.uniform(0.05, 0.95, size=ijk.shape)) / N
    up = rng.normal(size=(len(xp), 3))
    t = oracle.topo(3, [N + 1] * 3)
    info, ref = oracle.p2m(t, MP4, xp, up, [1, 1, 1])
    assert info == 0
    ref = oracle.unique_nodes(ref, [1, 1, 1], 3)
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, torch.float64), _dev(up, torch.float64)
    heads = np.concatenate([[0], np.cumsum(np.bincount(cells, minlength=N ** 3))])
    plan = core.Plan()
    fields = []
    for call in range(2):
        f = torch.empty(mesh.compact_shape(3), dtype=torch.float64, device="cuda")
        core.p2m(mesh, x, u, f, [0] * 3, MP4, F64, plan)
        assert _rel(f.cpu().numpy(), ref) < 1e-12, (case, call)
        fields.append(f)
        k, perm, start = plan.sort_view()
        assert np.array_equal(start.cpu().numpy(), heads), (case, call)
        assert np.array_equal(k.cpu().numpy(), cells), (case, call)
        assert np.array_equal(perm.cpu().numpy(), np.arange(len(xp))), (case, call)
    assert torch.equal(fields[0], fields[1])
    uniq = rng.normal(size=(N, N, N, 3))
    oref = bf.m2p_periodic(xp, uniq, N)
    mplan = core.Plan()
    for call in range(2):
        out = torch.empty((len(xp), 3), dtype=torch.float64, device="cuda")
        core.m2p(mesh, x, _dev(uniq, torch.float64), [0] * 3, out, MP4, F64, mplan)
        assert _rel(out.cpu().numpy(), oref) < 1e-12, (case, call)


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lo", [0, -2])
def test_binning_slot_table(core, oracle, kernel, prec, lo, monkeypatch):
    """Unsorted input for the tiled 3D deposit on a periodic mesh (compact
    layout, lo = 0, or the reference's stored box 1-g..Nm+g, lo = -2, whose
    ghost shell and adjoint read the table too) is binned into a slot table
    (up to eight particles per cell, in increasing index) instead of heads
    and a permutation.  One plan is fed, in shuffled order:
    an advected lattice (table); the same with 40 particles moved into one
    cell (the table overflows, the binning runs again with keys and slots, and
    the plan bins its next 16 inputs without a table); the lattice for those
    16 calls; the lattice again (table).  Every field must equal bitwise the
    field of the same particles binned with heads (PPMG_NO_TABLE=1), whose
    per-node addition order is the same, and the oracle's within tolerance;
    the adjoint (heads built from the table) must equal the heads binning's
    bitwise; sort_view must give numpy's stable argsort, keys and heads."""
    N, lda = 24, 3
    rng = np.random.default_rng(31 + kernel - lo)
    lattice = wl.advect(wl.jittered_lattice(N, seed=31), wl.fourier_modes(31), N)[0]
    g = [-lo] * 3
    crowded = lattice.copy()
    crowded[:40] = (np.array([5, 6, 7]) + rng.uniform(0.1, 0.9, size=(40, 3))) / N
    t = oracle.topo(3, [N + 1] * 3)
    mesh = core.Mesh(3, N + 1)
    plan = core.Plan()
    calls = [("lattice", True)] + [("crowded", True)] + [("lattice", False)] * 16 + [("lattice", True)]
    for i, (case, check) in enumerate(calls):
        base = lattice if case == "lattice" else crowded
        o = rng.permutation(len(base))
        xp = base[o]
        up = wl.strengths(i, len(xp), lda)
        x, u = _dev(xp, _pos_dtype(prec)), _dev(up, _val_dtype(prec))
        shape = mesh.compact_shape(lda) if lo == 0 else mesh.padded_shape(lda, g)
        f = torch.empty(shape, dtype=_val_dtype(prec), device="cuda")
        core.p2m(mesh, x, u, f, [lo] * 3, kernel, prec, plan)
        monkeypatch.setenv("PPMG_NO_TABLE", "1")
        href = torch.empty_like(f)
        hplan = core.Plan()
        core.p2m(mesh, x, u, href, [lo] * 3, kernel, prec, hplan)
        monkeypatch.delenv("PPMG_NO_TABLE")
        assert torch.equal(f, href), (case, i)
        if not check:
            continue
        if lo == 0:
            info, ref = oracle.p2m(t, kernel, xp, up, [1, 1, 1])
            ref = oracle.unique_nodes(ref, [1, 1, 1], 3)
        else:
            info, ref = oracle.p2m(t, kernel, xp, up, g)   # every stored node, ghosts raw
        assert info == 0
        assert _rel(f.cpu().numpy(), ref) < TOL[prec], (case, i)
        cot = _dev(rng.normal(size=f.shape), _val_dtype(prec))
        gu, gx = core.p2m_adjoint(mesh, x, u, cot, [lo] * 3, plan, kernel, prec)
        hu, hx = core.p2m_adjoint(mesh, x, u, cot, [lo] * 3, hplan, kernel, prec)
        assert torch.equal(gu, hu) and torch.equal(gx, hx), (case, i)
        # keys from the positions in the form's precision
        cells = ((xp.astype(np.float32) * np.float32(N)) if prec == F32 else xp * N).astype(np.int64)
        key = cells[:, 0] + N * (cells[:, 1] + N * cells[:, 2])
        k, perm, start = plan.sort_view()
        assert np.array_equal(k.cpu().numpy(), key), (case, i)
        assert np.array_equal(perm.cpu().numpy(), np.argsort(key, kind="stable")), (case, i)
        assert np.array_equal(start.cpu().numpy(),
                              np.concatenate([[0], np.cumsum(np.bincount(key, minlength=N ** 3))])), (case, i)


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("lo", [0, -2])
def test_m2p_slot_table(core, oracle, kernel, lo, monkeypatch):
    """The tiled 3D gather on a periodic mesh reads a slot-table binning of
    unsorted particles (any field layout), directly or left by
    ppmg_m2p_prebin.  One plan gathers at an advected lattice (table), at the
    same with 40 particles in one cell (the table overflows: heads), and at
    the lattice again after a prebinning; every result equals bitwise the
    gather of the same particles binned with heads (PPMG_NO_TABLE=1), since
    each particle's sum does not depend on the binning, and the oracle's
    within 1e-12."""
    N, lda = 24, 3
    rng = np.random.default_rng(61 + kernel - lo)
    lattice = wl.advect(wl.jittered_lattice(N, seed=61), wl.fourier_modes(61), N)[0]
    crowded = lattice.copy()
    crowded[:40] = (np.array([3, 9, 17]) + rng.uniform(0.1, 0.9, size=(40, 3))) / N
    uniq = rng.normal(size=(N, N, N, lda))
    mesh = core.Mesh(3, N + 1)
    if lo == 0:
        field = _dev(uniq, torch.float64)
    else:
        g = -lo
        idx = np.arange(-g, N + 1 + g) % N
        field = _dev(uniq[np.ix_(idx, idx, idx)], torch.float64)
    plan = core.Plan()
    for i, (case, prebin) in enumerate([("lattice", False), ("crowded", False), ("lattice", True)]):
        base = lattice if case == "lattice" else crowded
        xp = base[rng.permutation(len(base))]
        x = _dev(xp, torch.float64)
        out = torch.empty((len(xp), lda), dtype=torch.float64, device="cuda")
        if prebin:
            core.m2p_prebin(mesh, x, plan, kernel, F64)
        core.m2p(mesh, x, field, [lo] * 3, out, kernel, F64, plan)
        monkeypatch.setenv("PPMG_NO_TABLE", "1")
        ref = torch.empty_like(out)
        core.m2p(mesh, x, field, [lo] * 3, ref, kernel, F64)
        monkeypatch.delenv("PPMG_NO_TABLE")
        assert torch.equal(out, ref), (case, i)
        oref = bf.m2p_periodic(xp, uniq, N, kernel)
        assert _rel(out.cpu().numpy(), oref) < 1e-12, (case, i)


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C])
@pytest.mark.parametrize("lo", [0, -2])
def test_2d_slot_table(core, oracle, kernel, prec, lo, monkeypatch):
    """2D tiled deposit and gather on a periodic mesh, compact or the
    reference's stored box, with the slot-table binning of unsorted
    particles: a displaced lattice (table) and the same with 40 particles in
    one cell (the table overflows: heads).  Fields and gathered values equal
    bitwise those binned with heads (PPMG_NO_TABLE=1), the field the oracle's,
    the gather the brute-force sum; sort_view gives numpy's stable argsort and
    heads."""
    N, lda = 48, 3
    rng = np.random.default_rng(47 + kernel - lo + 10 * prec)
    base = (wl.jittered_lattice(N, seed=47, dim=2) + rng.uniform(-0.2, 0.2, size=(N * N, 2)) / N) % 1.0
    crowded = base.copy()
    crowded[:40] = (np.array([5, 9]) + rng.uniform(0.1, 0.9, size=(40, 2))) / N
    t = oracle.topo(2, [N + 1] * 2)
    mesh = core.Mesh(2, N + 1)
    g = [-lo] * 2
    vdt = _val_dtype(prec)
    uniq = rng.normal(size=(N, N, lda))
    for case, pts in (("lattice", base), ("crowded", crowded)):
        xp = pts[rng.permutation(len(pts))]
        up = rng.normal(size=(len(xp), lda))
        x, u = _dev(xp, torch.float64), _dev(up, vdt)
        shape = mesh.compact_shape(lda) if lo == 0 else mesh.padded_shape(lda, g)
        plan = core.Plan()
        f = torch.empty(shape, dtype=vdt, device="cuda")
        core.p2m(mesh, x, u, f, [lo] * 2, kernel, prec, plan)
        cells = (xp * N).astype(np.int64)
        key = cells[:, 0] + N * cells[:, 1]
        k, perm, start = plan.sort_view()
        assert np.array_equal(k.cpu().numpy(), key), case
        assert np.array_equal(perm.cpu().numpy(), np.argsort(key, kind="stable")), case
        assert np.array_equal(start.cpu().numpy(), np.concatenate([[0], np.cumsum(np.bincount(key, minlength=N * N))]))
        monkeypatch.setenv("PPMG_NO_TABLE", "1")
        href = torch.empty_like(f)
        core.p2m(mesh, x, u, href, [lo] * 2, kernel, prec)
        monkeypatch.delenv("PPMG_NO_TABLE")
        assert torch.equal(f, href), case
        info, ref = oracle.p2m(t, kernel, xp, up, [1, 1] if lo == 0 else g)
        assert info == 0
        if lo == 0:
            ref = oracle.unique_nodes(ref, [1, 1], 2)
        assert _rel(f.double().cpu().numpy(), ref) < TOL[prec], case
        field = _dev(uniq, vdt)
        out = torch.empty((len(xp), lda), dtype=vdt, device="cuda")
        core.m2p(mesh, x, field, [0] * 2, out, kernel, prec, core.Plan())
        monkeypatch.setenv("PPMG_NO_TABLE", "1")
        oref = torch.empty_like(out)
        core.m2p(mesh, x, field, [0] * 2, oref, kernel, prec)
        monkeypatch.delenv("PPMG_NO_TABLE")
        assert torch.equal(out, oref), case
        assert _rel(out.double().cpu().numpy(), bf.m2p_periodic(xp, uniq, N, kernel)) < TOL[prec], case


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
def test_m2p_lattice_without_binning(core, oracle, kernel, prec):
    """A plan whose last binning found a lattice in cell order (particle p in
    cell p, one per cell) gathers at the next such input without a binning
    pass, checking each particle's cell inside the gather; a particle out of
    its cell sends the call back to the binning.  One plan is fed: the
    lattice (binned; the lattice is recorded), the same (no pass), through
    ppmg_m2p_prebin (no pass), with particles 0 and 1 swapped (checked, then
    binned again; the plan then bins its next 16 inputs), the lattice, with
    particle 7 outside the domain (the general path), the lattice again.  Every output equals bitwise that of a fresh plan (a
    particle's sum does not depend on the binning) and the oracle's; after a
    checked call sort_view gives the identity heads and permutation."""
    N, lda = 20, 3
    rng = np.random.default_rng(83 + kernel + 10 * prec)
    lattice = wl.jittered_lattice(N, seed=83)
    cells = (lattice * N).astype(np.int64)
    assert np.array_equal(cells[:, 0] + N * (cells[:, 1] + N * cells[:, 2]), np.arange(N ** 3))
    uniq = rng.normal(size=(N, N, N, lda))
    vdt = _val_dtype(prec)
    mesh = core.Mesh(3, N + 1)
    field = _dev(uniq, vdt)
    swapped = lattice.copy()
    swapped[[0, 1]] = swapped[[1, 0]]
    outside = lattice.copy()
    outside[7] = [1.5, 0.5, 0.5]
    plan = core.Plan()
    for i, (case, xp, prebin) in enumerate([("lattice", lattice, False), ("lattice", lattice, False),
                                            ("prebin", lattice, True), ("swapped", swapped, False),
                                            ("lattice", lattice, False), ("outside", outside, False),
                                            ("lattice", lattice, False)]):
        x = _dev(xp, _pos_dtype(prec))
        out = torch.empty((len(xp), lda), dtype=vdt, device="cuda")
        if prebin:
            core.m2p_prebin(mesh, x, plan, kernel, prec)
        if case == "outside":
            # whatever a fresh plan does with a particle outside the domain
            # (an error, or the general path's values), this plan does too
            ref = torch.empty_like(out)
            try:
                core.m2p(mesh, x, field, [0] * 3, ref, kernel, prec)
            except RuntimeError as e:
                with pytest.raises(RuntimeError) as e2:
                    core.m2p(mesh, x, field, [0] * 3, out, kernel, prec, plan)
                assert str(e2.value) == str(e)
            else:
                core.m2p(mesh, x, field, [0] * 3, out, kernel, prec, plan)
                assert torch.equal(out, ref), (case, i)
            continue
        core.m2p(mesh, x, field, [0] * 3, out, kernel, prec, plan)
        ref = torch.empty_like(out)
        core.m2p(mesh, x, field, [0] * 3, ref, kernel, prec)
        assert torch.equal(out, ref), (case, i)
        oref = bf.m2p_periodic(xp, uniq, N, kernel)
        assert _rel(out.double().cpu().numpy(), oref) < TOL[prec], (case, i)
        if case == "lattice" and i >= 1:
            k, perm, start = plan.sort_view()
            assert np.array_equal(start.cpu().numpy(), np.arange(N ** 3 + 1)), (case, i)
            assert np.array_equal(perm.cpu().numpy(), np.arange(N ** 3)), (case, i)
            assert np.array_equal(k.cpu().numpy(), np.arange(N ** 3)), (case, i)


@pytest.mark.parametrize("dim,N,prec", [(3, 32, F64), (3, 300, F64), (2, 96, F64), (3, 32, F32), (2, 96, F32)])
def test_binning_crowded_radix(core, oracle, dim, N, prec, monkeypatch):
    """Crowded input is binned by a stable radix sort of (cell key, particle)
    pairs (8 bits a pass: two passes at 32^3 and 96^2 cells, four at 300^3),
    with the heads from the sorted runs; a plan takes that path after a
    binning that sampled many crowded cells, and PPMG_RADIX=1 forces it.
    One plan is fed a Gaussian blob (hundreds per cell) over a uniform
    background: shuffled (counting path), shuffled twice more and then in
    cell order under PPMG_RADIX=1 (radix; in cell order the keys pass finds
    no descent: identity permutation), and in cell order again (the
    cell-order path).  Every call: keys, permutation and heads equal numpy's
    stable argsort and bincount, and the field equals bitwise that of a
    fresh plan under PPMG_RADIX=0 (counting path), whose per-node addition
    order is the same; at 32^3 and 96^2 also the oracle's field.  FP32:
    positions and strengths in single precision (keys from the single
    precision product)."""
    rng = np.random.default_rng(900 + N + prec)
    nb, nu = (6000, 3000) if N < 200 else (60000, 30000)
    base = np.concatenate([rng.normal(0.4, 0.6 / N, size=(nb, dim)) % 1.0, rng.uniform(0, 1, size=(nu, dim))])
    odt = np.float32 if prec == F32 else np.float64
    base = base.astype(odt)
    cells = np.minimum((base * odt(N)).astype(np.int64), N - 1)
    key = cells[:, 0] + N * cells[:, 1] + (N * N * cells[:, 2] if dim == 3 else 0)
    ncells = N ** dim
    assert np.bincount(key).max() > 50
    mesh = core.Mesh(dim, N + 1)
    plan = core.Plan()
    t = oracle.topo(dim, [N + 1] * dim) if N < 200 else None
    for call, order in enumerate(["shuffled", "shuffled", "shuffled", "cell_sorted", "cell_sorted"]):
        o = np.argsort(key, kind="stable") if order == "cell_sorted" else rng.permutation(len(base))
        xp, k = base[o], key[o]
        up = rng.normal(size=(len(xp), 3)).astype(odt)
        x, u = _dev(xp, _pos_dtype(prec)), _dev(up, _val_dtype(prec))
        f = torch.empty(mesh.compact_shape(3), dtype=_val_dtype(prec), device="cuda")
        if call in (1, 2, 3):
            monkeypatch.setenv("PPMG_RADIX", "1")
        core.p2m(mesh, x, u, f, [0] * dim, MP4, prec, plan)
        monkeypatch.setenv("PPMG_RADIX", "0")
        kk, perm, start = plan.sort_view()
        assert np.array_equal(kk.cpu().numpy(), k), (order, call)
        assert np.array_equal(perm.cpu().numpy(), np.argsort(k, kind="stable")), (order, call)
        assert np.array_equal(start.cpu().numpy(),
                              np.concatenate([[0], np.cumsum(np.bincount(k, minlength=ncells))])), (order, call)
        ref = torch.empty_like(f)
        core.p2m(mesh, x, u, ref, [0] * dim, MP4, prec)
        monkeypatch.delenv("PPMG_RADIX")
        assert torch.equal(f, ref), (order, call)
        if t is not None:
            info, oref = oracle.p2m(t, MP4, xp, up, [1] * dim, dtype=odt)
            assert info == 0
            assert _rel(f.cpu().numpy(), oracle.unique_nodes(oref, [1] * dim, dim)) < TOL[prec], (order, call)


# ----------------------------------------------------------- variants --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lda", [1, 3, 5])
def test_p2m_variants(core, oracle, dim, kernel, prec, lda):
    N = 12 if dim == 3 else 20
    xp, up = _case(dim, N, 1500, 100 + dim * 10 + kernel + lda, lda, lo=-0.5, L=2.0)
    g = [2] * dim
    t = oracle.topo(dim, [N + 1] * dim, [-0.5] * dim, [1.5] * dim)
    odt = np.float32 if prec == F32 else np.float64
    info, ref = oracle.p2m(t, kernel, xp.astype(odt), up.astype(odt), g, dtype=odt)
    assert info == 0
    mesh = core.Mesh(dim, N + 1, [-0.5] * dim, [1.5] * dim)
    field = torch.zeros(mesh.padded_shape(lda, g), dtype=_val_dtype(prec), device="cuda")
    core.p2m(mesh, _dev(xp, _pos_dtype(prec)), _dev(up, _val_dtype(prec)), field, [-2] * dim, kernel, prec)
    assert _rel(field.cpu().numpy(), ref) < TOL[prec]


def test_p2m_compact_equals_padded_and_is_deterministic(core, monkeypatch):
    """Compact and padded (g = 1) layouts both run the tiled deposit; the
    padded one adds node N (equal to node 0 bitwise) and the raw ghost shell.
    Both against the general node gather, and bitwise reproducible."""
    N = 16
    xp, up = _case(3, N, 4000, 7, 3)
    mesh = core.Mesh(3, N + 1)
    padded = torch.zeros(mesh.padded_shape(3, [1, 1, 1]), dtype=torch.float64, device="cuda")
    compact = torch.zeros(mesh.compact_shape(3), dtype=torch.float64, device="cuda")
    x, u = _dev(xp, torch.float64), _dev(up, torch.float64)
    core.p2m(mesh, x, u, padded, [-1] * 3)
    core.p2m(mesh, x, u, compact, [0] * 3)
    again = torch.zeros_like(compact)
    core.p2m(mesh, x, u, again, [0] * 3)
    assert torch.equal(padded[1:1 + N, 1:1 + N, 1:1 + N], compact)
    assert torch.equal(padded[1:1 + N, 1:1 + N, 1 + N], padded[1:1 + N, 1:1 + N, 1])   # node N = node 0
    assert torch.equal(padded[1 + N, 1:2 + N, 1:2 + N], padded[1, 1:2 + N, 1:2 + N])
    assert torch.equal(compact, again)                      # bitwise reproducible
    ref = bf.p2m_periodic(xp, up, N)
    assert _rel(compact.cpu().numpy(), ref) < 1e-12
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.zeros_like(padded)
    core.p2m(mesh, x, u, gen, [-1] * 3)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(padded.cpu().numpy(), gen.cpu().numpy()) < 1e-13   # ghosts included


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lda", [1, 3, 4, 7])
def test_p2m_tiled_fast_path(core, oracle, kernel, prec, lda, monkeypatch):
    """Tiled 3D deposit (M'4 and BSP2, compact periodic layout) vs the oracle and vs the general path."""
    N = 40                                  # 13- and 15-node tiles leave a ragged last tile
    xp = wl.advect(wl.jittered_lattice(N, seed=8), wl.fourier_modes(8), N)[0]
    up = wl.strengths(8, len(xp), lda)
    t = oracle.topo(3, [N + 1] * 3)
    odt = np.float32 if prec == F32 else np.float64
    info, ref = oracle.p2m(t, kernel, xp.astype(odt), up.astype(odt), [1, 1, 1], dtype=odt)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, _pos_dtype(prec)), _dev(up, _val_dtype(prec))
    fast = torch.empty(mesh.compact_shape(lda), dtype=_val_dtype(prec), device="cuda")
    core.p2m(mesh, x, u, fast, [0] * 3, kernel, prec)
    assert _rel(fast.cpu().numpy(), oracle.unique_nodes(ref, [1, 1, 1], 3)) < TOL[prec]
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.empty_like(fast)
    core.p2m(mesh, x, u, gen, [0] * 3, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(fast.cpu().numpy(), gen.cpu().numpy()) < (1e-13 if prec == F64 else 1e-6)
    again = torch.empty_like(fast)
    core.p2m(mesh, x, u, again, [0] * 3, kernel, prec)
    assert torch.equal(fast, again)


@pytest.mark.parametrize("rows", [9, 13, 21, 29])
@pytest.mark.parametrize("N", [50, 64])
def test_p2m_tile_heights(core, rows, N, monkeypatch):
    """Every tile height of the tiled deposit (1 to 4 wavefronts sharing the
    LDS ring, the planes written out and cleared by all of them): the field
    equals the general path's to rounding, its unique nodes hold the total
    strength, and three launches agree bitwise.  N = 50 leaves ragged tiles
    in x and y for every height."""
    monkeypatch.setenv("PPMG_P2M_ROWS", str(rows))
    xp = wl.advect(wl.jittered_lattice(N, seed=rows), wl.fourier_modes(rows), N)[0]
    up = wl.strengths(rows, len(xp), 3)
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, torch.float64), _dev(up, torch.float64)
    runs = []
    for _ in range(3):
        f = torch.empty(mesh.compact_shape(3), dtype=torch.float64, device="cuda")
        core.p2m(mesh, x, u, f, [0] * 3, MP4, F64)
        runs.append(f)
    assert torch.equal(runs[0], runs[1]) and torch.equal(runs[0], runs[2])
    got = runs[0].cpu().numpy()
    assert np.allclose(got.reshape(-1, 3).sum(0), up.sum(0), rtol=0, atol=1e-10 * np.abs(up).sum())
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.empty_like(runs[0])
    core.p2m(mesh, x, u, gen, [0] * 3, MP4, F64)
    assert _rel(got, gen.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("rows", [9, 13, 29])
@pytest.mark.parametrize("lda", [1, 2, 4])
def test_p2m_tile_heights_components(core, rows, lda, monkeypatch):
    """Tile heights against component counts: the write-out hands each thread
    one (node, component) of one node row per pass, 64 to 256 threads over rows
    of 13 x lda values.  Ragged tiles (N = 50) against the general path."""
    monkeypatch.setenv("PPMG_P2M_ROWS", str(rows))
    N = 50
    xp = wl.advect(wl.jittered_lattice(N, seed=rows + lda), wl.fourier_modes(rows), N)[0]
    up = wl.strengths(rows + lda, len(xp), lda)
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, torch.float64), _dev(up, torch.float64)
    f = torch.empty(mesh.compact_shape(lda), dtype=torch.float64, device="cuda")
    core.p2m(mesh, x, u, f, [0] * 3, MP4, F64)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.empty_like(f)
    core.p2m(mesh, x, u, gen, [0] * 3, MP4, F64)
    assert _rel(f.cpu().numpy(), gen.cpu().numpy()) < 1e-13


# tiled against general on the same inputs: only the summation order differs
GEN_TOL = {F64: 1e-13, F32C: 1e-6, F32: 1e-5}


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("N", [5, 12, 40])
@pytest.mark.parametrize("g", [0, 1, 2])
def test_tiled_paths_padded_layout(core, N, g, prec, kernel, monkeypatch):
    """The tiled kernels on the reference's stored box (lo = -g, ext = N+1+2g)
    against the general paths: P2M on every stored node (node N equal to node
    0, ghost layers the raw unwrapped sums), and M2P of a random stored field
    whose node N, node 0 and ghost layers all differ, which the gather reads
    with ghost_get's node map.  N = 5 and 12 give tiles wider than the mesh."""
    rng = np.random.default_rng(1000 * N + g)
    lda = 2
    xp = rng.uniform(0.0, 1.0, size=(max(8 * N ** 3, 2000), 3))
    up = rng.normal(size=(len(xp), lda))
    mesh = core.Mesh(3, N + 1)
    lo = [-g] * 3
    vdt = _val_dtype(prec)
    x, u = _dev(xp, _pos_dtype(prec)), _dev(up, vdt)
    fast = torch.full(mesh.padded_shape(lda, [g] * 3), 7.0, dtype=vdt, device="cuda")
    core.p2m(mesh, x, u, fast, lo, kernel, prec)
    field = _dev(rng.normal(size=mesh.padded_shape(lda, [g] * 3)), vdt)
    out = torch.empty((len(xp), lda), dtype=vdt, device="cuda")
    core.m2p(mesh, x, field, lo, out, kernel, prec)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.full_like(fast, 7.0)
    core.p2m(mesh, x, u, gen, lo, kernel, prec)
    out_gen = torch.empty_like(out)
    core.m2p(mesh, x, field, lo, out_gen, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(fast.cpu().numpy(), gen.cpu().numpy()) < GEN_TOL[prec]
    s = slice(g, g + N)
    assert torch.equal(fast[s, s, g + N], fast[s, s, g])          # node N = node 0 (x)
    assert _rel(out.cpu().numpy(), out_gen.cpu().numpy()) < GEN_TOL[prec]


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("bc", [[1, 1, 1, 1, 2, 2], [2, 2, 1, 1, 1, 1], [2, 2, 2, 2, 2, 2]],
                         ids=["open_z", "open_x", "all_open"])
@pytest.mark.parametrize("N", [5, 12, 40])
def test_tiled_paths_open_axes(core, oracle, bc, N, prec, kernel, monkeypatch):
    """Meshes with non-periodic axes on the tiled kernels: on an open axis the
    tiles make the N+1 real nodes 0..N with no wrap, ghost nodes keep their raw
    sums, and the gather reads node v itself.  Against the general paths on
    every stored node (P2M with ghost sizes 0..2, M2P of a random stored field
    with ghost size 2), and against the oracle (FP32 against its float form).
    A tenth of the particles sit exactly on a face of an open axis (cell N
    holds the upper face)."""
    rng = np.random.default_rng(7 * N + sum(bc))
    lda = 3
    npart = max(6 * N ** 3, 3000)
    xp = rng.uniform(0.0, 1.0, size=(npart, 3))
    for d in range(3):
        if bc[2 * d] != 1:
            face = rng.uniform(size=npart) < 0.05
            xp[face, d] = 1.0
            face = rng.uniform(size=npart) < 0.05
            xp[face, d] = 0.0
    up = rng.normal(size=(npart, lda))
    vdt = _val_dtype(prec)
    odt = np.float32 if prec == F32 else np.float64
    mesh = core.Mesh(3, N + 1, bcdef=bc)
    x, u = _dev(xp, _pos_dtype(prec)), _dev(up, vdt)
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    for g in (0, 1, 2):
        lo = [-g] * 3
        fast = torch.full(mesh.padded_shape(lda, [g] * 3), 7.0, dtype=vdt, device="cuda")
        core.p2m(mesh, x, u, fast, lo, kernel, prec)
        monkeypatch.setenv("PPMG_GENERAL", "1")
        gen = torch.full_like(fast, 7.0)
        core.p2m(mesh, x, u, gen, lo, kernel, prec)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(fast.cpu().numpy(), gen.cpu().numpy()) < GEN_TOL[prec], g
        if g == 2:
            info, ref = oracle.p2m(t, kernel, xp.astype(odt), up.astype(odt), [2, 2, 2], dtype=odt)
            assert info == 0
            assert _rel(fast.cpu().numpy(), ref) < TOL[prec]
    fr = rng.normal(size=mesh.padded_shape(lda, [2] * 3))
    field = _dev(fr, vdt)
    out = torch.empty((npart, lda), dtype=vdt, device="cuda")
    core.m2p(mesh, x, field, [-2] * 3, out, kernel, prec)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    out_gen = torch.empty_like(out)
    core.m2p(mesh, x, field, [-2] * 3, out_gen, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(out.cpu().numpy(), out_gen.cpu().numpy()) < GEN_TOL[prec]
    # the oracle runs ghost_get first: beside an open axis it fills only the
    # nodes whose open coordinates are real nodes, the rest read as stored
    info, qref = oracle.m2p(t, kernel, xp.astype(odt), fr.astype(odt), [2] * 3, dtype=odt)
    assert info == 0
    assert _rel(out.cpu().numpy(), qref) < TOL[prec]
    if any(b == 1 for b in bc):
        # compact periodic axes beside boxed open ones: the tiled gather
        # against the general one on the same layout
        lo = [0 if bc[2 * d] == 1 else -2 for d in range(3)]
        shape = tuple(reversed([N if bc[2 * d] == 1 else N + 5 for d in range(3)])) + (lda,)
        fm = _dev(rng.normal(size=shape), vdt)
        core.m2p(mesh, x, fm, lo, out, kernel, prec)
        pm = torch.full(shape, 7.0, dtype=vdt, device="cuda")
        core.p2m(mesh, x, u, pm, lo, kernel, prec)
        monkeypatch.setenv("PPMG_GENERAL", "1")
        core.m2p(mesh, x, fm, lo, out_gen, kernel, prec)
        pg = torch.full_like(pm, 7.0)
        core.p2m(mesh, x, u, pg, lo, kernel, prec)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(out.cpu().numpy(), out_gen.cpu().numpy()) < GEN_TOL[prec]
        assert _rel(pm.cpu().numpy(), pg.cpu().numpy()) < GEN_TOL[prec]


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("bc", [[1, 1, 1, 1], [1, 1, 2, 2], [2, 2, 2, 2]], ids=["periodic", "open_y", "all_open"])
@pytest.mark.parametrize("N", [5, 12, 40, 301])
def test_p2m2d_tiled(core, oracle, bc, N, prec, kernel, monkeypatch):
    """The 2D tiled deposit (13- or 15-node strips marching four cell rows a
    step) on the compact layout and on the reference's stored box (ghost
    sizes 0..2), periodic and open axes: every stored node against the
    general path, the ghost-size-2 box against the oracle, and two launches
    bitwise equal.  N = 40 leaves a ragged last strip; N = 301 splits the rows
    into several chunks.  A tenth of the particles sit on a face of an open
    axis."""
    rng = np.random.default_rng(11 * N + sum(bc) + kernel)
    lda = 3
    npart = max(3 * N * N, 2000)
    xp = rng.uniform(0.0, 1.0, size=(npart, 2))
    for d in range(2):
        if bc[2 * d] != 1:
            xp[rng.uniform(size=npart) < 0.05, d] = 1.0
            xp[rng.uniform(size=npart) < 0.05, d] = 0.0
    up = rng.normal(size=(npart, lda))
    vdt = _val_dtype(prec)
    odt = np.float32 if prec == F32 else np.float64
    mesh = core.Mesh(2, N + 1, bcdef=bc)
    x, u = _dev(xp, _pos_dtype(prec)), _dev(up, vdt)
    layouts = [(g, [-g] * 2, mesh.padded_shape(lda, [g] * 2)) for g in (0, 1, 2)]
    if all(b == 1 for b in bc):
        layouts.append(("compact", [0, 0], mesh.compact_shape(lda)))
    for g, lo, shape in layouts:
        fast = torch.full(shape, 7.0, dtype=vdt, device="cuda")
        core.p2m(mesh, x, u, fast, lo, kernel, prec)
        again = torch.full(shape, 7.0, dtype=vdt, device="cuda")
        core.p2m(mesh, x, u, again, lo, kernel, prec)
        assert torch.equal(fast, again), g
        monkeypatch.setenv("PPMG_GENERAL", "1")
        gen = torch.full_like(fast, 7.0)
        core.p2m(mesh, x, u, gen, lo, kernel, prec)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(fast.cpu().numpy(), gen.cpu().numpy()) < GEN_TOL[prec], g
        if g == 2:
            t = oracle.topo(2, [N + 1] * 2, bcdef=bc)
            info, ref = oracle.p2m(t, kernel, xp.astype(odt), up.astype(odt), [2, 2], dtype=odt)
            assert info == 0
            assert _rel(fast.cpu().numpy(), ref) < TOL[prec]


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("bc", [[1, 1, 1, 1], [1, 1, 2, 2], [2, 2, 2, 2]], ids=["periodic", "open_y", "all_open"])
@pytest.mark.parametrize("N", [5, 12, 40, 301])
def test_m2p2d_tiled(core, oracle, bc, N, prec, kernel, monkeypatch):
    """The 2D tiled gather on random stored fields whose ghosts, node N and
    node 0 all differ: the reference's stored box (ghost size 2) against the
    oracle (which runs ghost_get first) and the general path; on a periodic
    mesh also ghost sizes 0 and 1 and the compact layout against the general
    path; particles in random order, a tenth of them on faces of open axes.
    Beside an open axis a periodic box takes the general path (same result)."""
    rng = np.random.default_rng(13 * N + sum(bc) + kernel)
    lda = 3
    npart = max(3 * N * N, 2000)
    xp = rng.uniform(0.0, 1.0, size=(npart, 2))
    for d in range(2):
        if bc[2 * d] != 1:
            xp[rng.uniform(size=npart) < 0.05, d] = 1.0
            xp[rng.uniform(size=npart) < 0.05, d] = 0.0
    vdt = _val_dtype(prec)
    odt = np.float32 if prec == F32 else np.float64
    mesh = core.Mesh(2, N + 1, bcdef=bc)
    x = _dev(xp, _pos_dtype(prec))
    layouts = [(2, [-2, -2], mesh.padded_shape(lda, [2, 2]))]
    if all(b == 1 for b in bc):
        layouts += [(g, [-g] * 2, mesh.padded_shape(lda, [g] * 2)) for g in (0, 1)]
        layouts.append(("compact", [0, 0], mesh.compact_shape(lda)))
    for g, lo, shape in layouts:
        fr = rng.normal(size=shape)
        f = _dev(fr, vdt)
        out = torch.full((npart, lda), 7.0, dtype=vdt, device="cuda")
        core.m2p(mesh, x, f, lo, out, kernel, prec)
        monkeypatch.setenv("PPMG_GENERAL", "1")
        gen = torch.full_like(out, 7.0)
        core.m2p(mesh, x, f, lo, gen, kernel, prec)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(out.cpu().numpy(), gen.cpu().numpy()) < GEN_TOL[prec], g
        if g == 2:
            t = oracle.topo(2, [N + 1] * 2, bcdef=bc)
            info, ref = oracle.m2p(t, kernel, xp.astype(odt), fr.astype(odt), [2, 2], dtype=odt)
            assert info == 0
            assert _rel(out.cpu().numpy(), ref) < TOL[prec]


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C])
@pytest.mark.parametrize("spread", ["one_cell", "blob", "mixed"])
def test_tiled_crowded_cells(core, oracle, kernel, prec, spread, monkeypatch):
    """Cells holding tens to hundreds of particles, where the tile kernels
    leave the round structure (the deposit's one-at-a-time path with its
    particles loaded 64 at a time, the gather's one-particle-per-lane pass):
    P2M on every node and M2P at every particle against the general path and
    the oracle, and two launches bitwise equal.  one_cell: 300 particles in
    one cell beside a sparse uniform background; blob: a Gaussian blob a few
    cells wide; mixed: several crowded cells in one 16-cell row plus light
    ones."""
    N = 24
    rng = np.random.default_rng({"one_cell": 1, "blob": 2, "mixed": 3}[spread] + 10 * kernel)
    h = 1.0 / N
    bg = rng.uniform(0.0, 1.0, size=(3000, 3))
    if spread == "one_cell":
        dense = (np.array([9, 10, 11]) + rng.uniform(0.05, 0.95, size=(300, 3))) * h
    elif spread == "blob":
        dense = np.mod(0.4 + 0.03 * rng.normal(size=(20000, 3)), 1.0)
    else:
        cells = np.array([[2 + 2 * k, 7, 13] for k in range(6)] + [[3, 8, 13], [5, 6, 12]])
        counts = rng.integers(20, 140, size=len(cells))
        dense = np.concatenate([(c + rng.uniform(0.02, 0.98, size=(n, 3))) * h for c, n in zip(cells, counts)])
    xp = np.concatenate([bg, dense])[rng.permutation(len(bg) + len(dense))]
    up = rng.normal(size=(len(xp), 3))
    vdt = _val_dtype(prec)
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, torch.float64), _dev(up, vdt)
    f1 = torch.empty(mesh.compact_shape(3), dtype=vdt, device="cuda")
    core.p2m(mesh, x, u, f1, [0] * 3, kernel, prec)
    f2 = torch.empty_like(f1)
    core.p2m(mesh, x, u, f2, [0] * 3, kernel, prec)
    assert torch.equal(f1, f2)
    field = _dev(rng.normal(size=mesh.compact_shape(3)), vdt)
    o1 = torch.empty((len(xp), 3), dtype=vdt, device="cuda")
    core.m2p(mesh, x, field, [0] * 3, o1, kernel, prec)
    o2 = torch.empty_like(o1)
    core.m2p(mesh, x, field, [0] * 3, o2, kernel, prec)
    assert torch.equal(o1, o2)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    fg = torch.empty_like(f1)
    core.p2m(mesh, x, u, fg, [0] * 3, kernel, prec)
    og = torch.empty_like(o1)
    core.m2p(mesh, x, field, [0] * 3, og, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(f1.cpu().numpy(), fg.cpu().numpy()) < GEN_TOL[prec]
    assert _rel(o1.cpu().numpy(), og.cpu().numpy()) < GEN_TOL[prec]
    t = oracle.topo(3, [N + 1] * 3)
    info, ref = oracle.p2m(t, kernel, xp, up, [1, 1, 1])
    assert info == 0
    assert _rel(f1.double().cpu().numpy(), oracle.unique_nodes(ref, [1, 1, 1], 3)) < TOL[prec]
    qref = bf.m2p_periodic(xp, field.double().cpu().numpy(), N, kernel)
    assert _rel(o1.double().cpu().numpy(), qref) < TOL[prec]


@pytest.mark.parametrize("bc", [[1] * 6, [2] * 6, [1, 1, 1, 1, 2, 2]], ids=["periodic", "all_open", "open_z"])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_smallest_meshes(core, oracle, bc, N, monkeypatch):
    """Meshes of 1-4 cells per axis, where every stencil wraps onto itself
    (ppm_topo_mkfield.f:421 allows Nm >= 2, one cell).  N = 4 is the
    smallest mesh the tiled kernels take, N = 1-3 run the general paths.  P2M
    and M2P on the reference's stored box (ghost size 2) against the oracle
    (at N = 1 the fold's one-period images leave the nodes two periods away
    in the ghost layers, as the reference's block lists do); on a periodic
    mesh also the compact layout against the numpy brute force; at N = 4 the
    tiled kernels against the general ones."""
    rng = np.random.default_rng(40 + 3 * N + sum(bc))
    lda = 3
    xp = rng.uniform(0.0, 1.0, size=(600, 3))
    for d in range(3):
        if bc[2 * d] != 1:
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 1.0
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 0.0
    up = rng.normal(size=(len(xp), lda))
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    mesh = core.Mesh(3, N + 1, bcdef=bc)
    x, u = _dev(xp, torch.float64), _dev(up, torch.float64)
    g = [2] * 3
    info, ref = oracle.p2m(t, MP4, xp, up, g)
    assert info == 0
    field = torch.full(mesh.padded_shape(lda, g), 7.0, dtype=torch.float64, device="cuda")
    core.p2m(mesh, x, u, field, [-2] * 3)
    assert _rel(field.cpu().numpy(), ref) < 1e-12
    fr = rng.normal(size=oracle.field_shape(t, lda, g))
    info, qref = oracle.m2p(t, MP4, xp, fr, g)
    assert info == 0
    q = torch.empty((len(xp), lda), dtype=torch.float64, device="cuda")
    core.m2p(mesh, x, _dev(fr, torch.float64), [-2] * 3, q)
    assert _rel(q.cpu().numpy(), qref) < 1e-12
    # the adjoint of the stored-box P2M: <P2M(u), g> = <u, P2M^T g> for a random cotangent g
    gco = rng.normal(size=oracle.field_shape(t, lda, g))
    plan = core.p2m(mesh, x, u, field, [-2] * 3)
    gu, _ = core.p2m_adjoint(mesh, x, u, _dev(gco, torch.float64), [-2] * 3, plan, want_x=False)
    lhs = float((field.cpu().numpy() * gco).sum())
    rhs = float((up * gu.cpu().numpy()).sum())
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1.0)
    if bc == [1] * 6:
        comp = torch.empty(mesh.compact_shape(lda), dtype=torch.float64, device="cuda")
        core.p2m(mesh, x, u, comp, [0] * 3)
        assert _rel(comp.cpu().numpy(), bf.p2m_periodic(xp, up, N)) < 1e-12
        uniq = rng.normal(size=(N, N, N, lda))
        core.m2p(mesh, x, _dev(uniq, torch.float64), [0] * 3, q)
        assert _rel(q.cpu().numpy(), bf.m2p_periodic(xp, uniq, N)) < 1e-12
    if N == 4:
        monkeypatch.setenv("PPMG_GENERAL", "1")
        gen = torch.full_like(field, 7.0)
        core.p2m(mesh, x, u, gen, [-2] * 3)
        qg = torch.empty_like(q)
        core.m2p(mesh, x, _dev(fr, torch.float64), [-2] * 3, qg)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(field.cpu().numpy(), gen.cpu().numpy()) < 1e-13
        q2 = torch.empty_like(q)
        core.m2p(mesh, x, _dev(fr, torch.float64), [-2] * 3, q2)
        assert _rel(q2.cpu().numpy(), qg.cpu().numpy()) < 1e-13


def test_empty_inputs(core, oracle):
    """No particles: every entry point returns without error and matches the
    oracle's empty results (P2M and remesh leave a zero field, the ranking
    gives empty boxes, create_part of a zero field selects nothing)."""
    N = 8
    mesh = core.Mesh(3, N + 1)
    t = oracle.topo(3, [N + 1] * 3)
    x = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    u = torch.zeros((0, 2), dtype=torch.float64, device="cuda")
    field = torch.full(mesh.compact_shape(2), 3.0, dtype=torch.float64, device="cuda")
    plan = core.p2m(mesh, x, u, field, [0] * 3)
    assert torch.all(field == 0)
    q = torch.zeros((0, 2), dtype=torch.float64, device="cuda")
    core.m2p(mesh, x, field, [0] * 3, q)
    gu, gx = core.p2m_adjoint(mesh, x, u, field, [0] * 3, plan)
    assert gu.shape == (0, 2) and gx.shape == (0, 3)
    lp, lh = core.util_rank(x, [0.0] * 3, [1.0] * 3, [4, 4, 4], [0] * 6)
    lp_ref, lh_ref, ic, _, _ = oracle.rank(np.zeros((0, 3)), [0.0] * 3, [1.0] * 3, [4, 4, 4], [0] * 6)
    assert ic == 0 and lp.numel() == 0 and np.array_equal(lh.cpu().numpy(), lh_ref)
    npbx = core.util_sort(x, [0.0] * 3, [1.0] * 3, [4, 4, 4])
    assert torch.all(npbx == 0)
    core.impose_part_bc(mesh, x)
    w = core.comp_weights(mesh, x)
    assert all(wd.shape == (0, 4) for wd in w)
    g = [2] * 3
    info, fref = oracle.remesh(t, MP4, np.zeros((0, 3)), np.zeros((0, 2)), g, [np.zeros((0, 4))] * 3)
    assert info == 0
    rf = torch.full(mesh.padded_shape(2, g), 5.0, dtype=torch.float64, device="cuda")
    core.remesh(mesh, x, u, w, rf, [-2] * 3)
    assert np.array_equal(rf.cpu().numpy(), fref)
    zero = torch.zeros(mesh.padded_shape(2, [1] * 3), dtype=torch.float64, device="cuda")
    xo, uo = core.create_part(mesh, zero, [-1] * 3, [0.5, 1.0])
    assert xo.shape == (0, 3) and uo.shape == (0, 2)


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C])
def test_tiled_paths_noncubic_box(core, oracle, prec, kernel, monkeypatch):
    """The three tiled kernels on a non-cubic mesh over a non-unit box: 40 x 29
    x 18 cells on [0,2] x [-1,0.5] x [0.25,1] (a different cell count, spacing
    and origin per axis, ragged tiles in x and y, a z extent under one chunk).
    P2M and M2P against the oracle, the adjoint against the oracle and the
    general per-particle path."""
    Ns = (40, 29, 18)
    mn, mx = (0.0, -1.0, 0.25), (2.0, 0.5, 1.0)
    lda = 3
    rng = np.random.default_rng(77)
    def box(n):
        return np.stack([mn[d] + (mx[d] - mn[d]) * rng.uniform(0.0, 1.0, n) for d in range(3)], 1)
    xp, xq = box(12000), box(5000)
    up = rng.normal(size=(len(xp), lda))
    t = oracle.topo(3, [n + 1 for n in Ns], mn, mx)
    info, ref = oracle.p2m(t, kernel, xp, up, [1, 1, 1])
    assert info == 0
    info, uref = oracle.m2p(t, kernel, xq, ref, [1, 1, 1])
    assert info == 0
    mesh = core.Mesh(3, [n + 1 for n in Ns], mn, mx)
    vdt = _val_dtype(prec)
    tol = 1e-12 if prec == F64 else 1e-6
    x, u = _dev(xp, _pos_dtype(prec)), _dev(up, vdt)
    field = torch.empty(mesh.compact_shape(lda), dtype=vdt, device="cuda")
    assert field.shape == (18, 29, 40, lda)
    plan = core.p2m(mesh, x, u, field, [0] * 3, kernel, prec)
    assert _rel(field.double().cpu().numpy(), oracle.unique_nodes(ref, [1, 1, 1], 3)) < tol
    out = torch.empty((len(xq), lda), dtype=vdt, device="cuda")
    core.m2p(mesh, _dev(xq, _pos_dtype(prec)), field, [0] * 3, out, kernel, prec)
    assert _rel(out.double().cpu().numpy(), uref) < tol
    gcot = rng.normal(size=field.shape)
    info, gu_ref, gx_ref = oracle.p2m_adjoint(t, kernel, xp, up, gcot)
    assert info == 0
    gc = _dev(gcot, vdt)
    gu, gx = core.p2m_adjoint(mesh, x, u, gc, [0] * 3, plan, kernel, prec)
    tol_adj = 1e-12 if prec == F64 else 2e-6
    assert _rel(gu.double().cpu().numpy(), gu_ref) < tol_adj
    assert _rel(gx.double().cpu().numpy(), gx_ref) < tol_adj
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gu2, gx2 = core.p2m_adjoint(mesh, x, u, gc, [0] * 3, plan, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(gu.double().cpu().numpy(), gu2.double().cpu().numpy()) < (1e-13 if prec == F64 else 1e-6)
    assert _rel(gx.double().cpu().numpy(), gx2.double().cpu().numpy()) < (1e-13 if prec == F64 else 1e-6)


def test_p2m_nonperiodic_axes(core, oracle):
    N = 10
    xp, up = _case(3, N, 800, 11, 2)
    bc = [2, 2, 1, 1, 2, 2]
    g = [2, 1, 2]
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    info, ref = oracle.p2m(t, MP4, xp, up, g)
    assert info == 0
    mesh = core.Mesh(3, N + 1, bcdef=bc)
    field = torch.zeros(mesh.padded_shape(2, g), dtype=torch.float64, device="cuda")
    core.p2m(mesh, _dev(xp, torch.float64), _dev(up, torch.float64), field, [-2, -1, -2])
    assert _rel(field.cpu().numpy(), ref) < 1e-12


def test_p2m_clustered_occupancy(core, oracle):
    """Gaussian blobs: cells with tens of particles; permutation stays stable."""
    N = 16
    xp = wl.gaussian_blobs(20000, seed=4)
    up = wl.strengths(4, len(xp), 3)
    t = oracle.topo(3, [N + 1] * 3)
    info, ref = oracle.p2m(t, MP4, xp, up, [1, 1, 1])
    mesh = core.Mesh(3, N + 1)
    field = torch.zeros(mesh.padded_shape(3, [1] * 3), dtype=torch.float64, device="cuda")
    plan = core.p2m(mesh, _dev(xp, torch.float64), _dev(up, torch.float64), field, [-1] * 3)
    assert _rel(field.cpu().numpy(), ref) < 1e-12
    key, perm, _ = plan.sort_view()
    k = key.cpu().numpy()
    assert np.bincount(k).max() > 50
    assert np.array_equal(perm.cpu().numpy(), np.argsort(k, kind="stable"))


def test_p2m_unassigned_particle_is_fatal_405(core):
    from ppm_legacy_amd import _lib
    mesh = core.Mesh(3, 9)
    xp = torch.tensor([[0.5, 0.5, 0.5], [0.5, 1.0, 0.5]], dtype=torch.float64, device="cuda")
    up = torch.ones((2, 1), dtype=torch.float64, device="cuda")
    field = torch.full(mesh.compact_shape(1), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.PpmgError) as e:
        core.p2m(mesh, xp, up, field, [0, 0, 0])
    assert e.value.info == -1 and e.value.code == _lib.ERR_PART_UNASS
    assert torch.all(field == 7.0)          # field untouched, like the reference's early return


def test_p2m_empty(core):
    mesh = core.Mesh(3, 9)
    xp = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    up = torch.zeros((0, 2), dtype=torch.float64, device="cuda")
    field = torch.full(mesh.compact_shape(2), 3.0, dtype=torch.float64, device="cuda")
    core.p2m(mesh, xp, up, field, [0, 0, 0])
    assert torch.all(field == 0)


# ----------------------------------------------------------------- M2P --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lda", [1, 3, 6])
def test_m2p_variants(core, oracle, dim, kernel, prec, lda):
    N = 12 if dim == 3 else 24
    g = [2] * dim
    rng = np.random.default_rng(200 + dim + kernel + lda)
    t = oracle.topo(dim, [N + 1] * dim)
    f = rng.normal(size=oracle.field_shape(t, lda, g))        # ghosts and node N differ from node 0
    xp = rng.uniform(0, 1, size=(2000, dim))
    odt = np.float32 if prec == F32 else np.float64
    info, ref = oracle.m2p(t, kernel, xp.astype(odt), f.astype(odt), g, dtype=odt)
    assert info == 0
    mesh = core.Mesh(dim, N + 1)
    up = torch.empty((2000, lda), dtype=_val_dtype(prec), device="cuda")
    core.m2p(mesh, _dev(xp, _pos_dtype(prec)), _dev(f, _val_dtype(prec)), [-2] * dim, up, kernel, prec)
    assert _rel(up.cpu().numpy(), ref) < TOL[prec]


@pytest.mark.parametrize("dim,bc", [(2, [1, 1, 2, 2]), (2, [2, 2, 1, 1]), (3, [2, 2, 1, 1, 1, 1]),
                                    (3, [1, 1, 2, 2, 2, 2])])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("gs", [2, 3])
def test_p2m_m2p_mixed_axes_vs_oracle(core, oracle, dim, bc, kernel, gs):
    """P2M, and M2P from a random stored box whose ghosts all differ, on a
    mesh with periodic and open axes, against the oracle.  ghost_get maps a
    node only when its open coordinates are real nodes 0..N, so ghost nodes
    beside an open face are read as stored on every axis.  A tenth of the
    particles sit on the open faces, where the stencils reach those nodes."""
    N = 9
    g = [gs] * dim
    rng = np.random.default_rng(60 + dim + kernel + sum(bc) + gs)
    xp = rng.uniform(0, 1, size=(3000, dim))
    for d in range(dim):
        if bc[2 * d] != 1:
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 1.0
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 0.0
    t = oracle.topo(dim, [N + 1] * dim, bcdef=bc)
    f = rng.normal(size=oracle.field_shape(t, 2, g))
    info, ref = oracle.m2p(t, kernel, xp, f, g)
    assert info == 0
    mesh = core.Mesh(dim, N + 1, bcdef=bc)
    x = _dev(xp, torch.float64)
    up = torch.empty((len(xp), 2), dtype=torch.float64, device="cuda")
    core.m2p(mesh, x, _dev(f, torch.float64), [-gs] * dim, up, kernel)
    assert _rel(up.cpu().numpy(), ref) < 1e-12
    info, fref = oracle.p2m(t, kernel, xp, up.cpu().numpy(), g)
    assert info == 0
    field = torch.full(fref.shape, 7.0, dtype=torch.float64, device="cuda")
    core.p2m(mesh, x, up, field, [-gs] * dim, kernel)
    assert _rel(field.cpu().numpy(), fref) < 1e-12


def test_m2p_stencil_outside_stored_field_is_9101(core):
    """A particle on the upper face of an open axis lies in cell N, and its
    M'4 stencil reaches node N+2, which a box with ghost size 1 does not
    store: the reference would read outside field_up, the library counts
    the read on the device and returns an error.  Interior particles alone
    read inside the box and succeed."""
    from ppm_legacy_amd import _lib
    N = 8
    mesh = core.Mesh(3, N + 1, bcdef=[1, 1, 1, 1, 2, 2])
    field = torch.ones(mesh.padded_shape(1, [1] * 3), dtype=torch.float64, device="cuda")
    inner = torch.tensor([[0.5, 0.5, 0.5], [0.1, 0.9, 0.45]], dtype=torch.float64, device="cuda")
    up = torch.empty((2, 1), dtype=torch.float64, device="cuda")
    core.m2p(mesh, inner, field, [-1] * 3, up)
    assert torch.allclose(up, torch.ones_like(up), rtol=0, atol=1e-14)
    face = torch.tensor([[0.5, 0.5, 0.5], [0.5, 0.5, 1.0]], dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.PpmgError) as e:
        core.m2p(mesh, face, field, [-1] * 3, up)
    assert e.value.code == _lib.ERR_FIELD_BOUNDS


def test_m2p_compact_layout(core):
    N = 16
    rng = np.random.default_rng(9)
    uniq = rng.normal(size=(N, N, N, 3))
    xp = rng.uniform(0, 1, size=(3000, 3))
    mesh = core.Mesh(3, N + 1)
    up = torch.empty((3000, 3), dtype=torch.float64, device="cuda")
    core.m2p(mesh, _dev(xp, torch.float64), _dev(uniq, torch.float64), [0, 0, 0], up)
    assert _rel(up.cpu().numpy(), bf.m2p_periodic(xp, uniq, N)) < 1e-12


@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lda", [1, 3, 4, 7])
def test_m2p_tiled_fast_path(core, oracle, kernel, prec, lda, monkeypatch):
    """Tiled 3D gather (M'4 and BSP2, compact periodic layout, particles in
    random order) vs the brute-force sum and the general path."""
    N = 40
    rng = np.random.default_rng(31 + lda)
    uniq = rng.normal(size=(N, N, N, lda))
    xp = rng.uniform(0, 1, size=(30000, 3))
    mesh = core.Mesh(3, N + 1)
    x = _dev(xp, _pos_dtype(prec))
    f = _dev(uniq, _val_dtype(prec))
    fast = torch.empty((len(xp), lda), dtype=_val_dtype(prec), device="cuda")
    core.m2p(mesh, x, f, [0] * 3, fast, kernel, prec)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gen = torch.empty_like(fast)
    core.m2p(mesh, x, f, [0] * 3, gen, kernel, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    tol = 1e-13 if prec == F64 else 1e-6
    assert _rel(fast.cpu().numpy(), gen.cpu().numpy()) < tol
    ref = bf.m2p_periodic(xp.astype(np.float32).astype(np.float64) if prec == F32 else xp, uniq, N, kernel)
    assert _rel(fast.cpu().numpy(), ref) < (1e-12 if prec == F64 else 1e-5)


def test_m2p_tiled_falls_back_for_particles_outside(core):
    N = 16
    rng = np.random.default_rng(2)
    uniq = rng.normal(size=(N, N, N, 2))
    xp = rng.uniform(0, 1, size=(500, 3))
    xp[7] = [1.0 + 0.25 / N, 0.5, 0.5]                  # past max_phys: read through the periodic map
    mesh = core.Mesh(3, N + 1)
    up = torch.empty((500, 2), dtype=torch.float64, device="cuda")
    core.m2p(mesh, _dev(xp, torch.float64), _dev(uniq, torch.float64), [0] * 3, up)
    assert _rel(up.cpu().numpy(), bf.m2p_periodic(xp, uniq, N)) < 1e-12


# ------------------------------------------------------------- adjoint --
@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C])
def test_adjoint_vs_oracle(core, oracle, kernel, prec):
    N = 16
    xp, up = _case(3, N, 3000, 300 + kernel, 2)
    gcot = np.random.default_rng(5).normal(size=(N, N, N, 2))
    t = oracle.topo(3, [N + 1] * 3)
    info, gu_ref, gx_ref = oracle.p2m_adjoint(t, kernel, xp, up, gcot)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    x, u = _dev(xp, torch.float64), _dev(up, _val_dtype(prec))
    field = torch.empty(mesh.compact_shape(2), dtype=_val_dtype(prec), device="cuda")
    plan = core.p2m(mesh, x, u, field, [0] * 3, kernel, prec)
    gu, gx = core.p2m_adjoint(mesh, x, u, _dev(gcot, _val_dtype(prec)), [0] * 3, plan, kernel, prec)
    tol = 1e-12 if prec == F64 else 2e-6
    assert _rel(gu.cpu().numpy(), gu_ref) < tol
    assert _rel(gx.cpu().numpy(), gx_ref) < tol


@pytest.mark.parametrize("prec", [F64, F32C, F32])
@pytest.mark.parametrize("lda", [1, 3, 5])
def test_adjoint_tiled_vs_oracle_and_general(core, oracle, prec, lda, monkeypatch):
    """Tiled 3D M'4 adjoint (compact periodic cotangent, several components,
    ragged tiles, particles in random order) vs the oracle and vs the general
    per-particle path."""
    N = 40
    rng = np.random.default_rng(90 + lda)
    xp = rng.uniform(0.0, 1.0, size=(6000, 3))
    up = rng.normal(size=(len(xp), lda))
    gcot = rng.normal(size=(N, N, N, lda))
    t = oracle.topo(3, [N + 1] * 3)
    info, gu_ref, gx_ref = oracle.p2m_adjoint(t, MP4, xp, up, gcot)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    vdt = _val_dtype(prec)
    x, u, gc = _dev(xp, _pos_dtype(prec)), _dev(up, vdt), _dev(gcot, vdt)
    field = torch.empty(mesh.compact_shape(lda), dtype=vdt, device="cuda")
    plan = core.p2m(mesh, x, u, field, [0] * 3, MP4, prec)
    gu, gx = core.p2m_adjoint(mesh, x, u, gc, [0] * 3, plan, MP4, prec)
    tol = {F64: 1e-12, F32C: 2e-6, F32: 1e-5}[prec]
    assert _rel(gu.cpu().numpy(), gu_ref) < tol
    assert _rel(gx.cpu().numpy(), gx_ref) < tol
    monkeypatch.setenv("PPMG_GENERAL", "1")
    gu2, gx2 = core.p2m_adjoint(mesh, x, u, gc, [0] * 3, plan, MP4, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    assert _rel(gu.cpu().numpy(), gu2.cpu().numpy()) < (1e-13 if prec == F64 else 1e-6)
    assert _rel(gx.cpu().numpy(), gx2.cpu().numpy()) < (1e-13 if prec == F64 else 1e-6)


def test_autograd_matches_fd_and_oracle(core, oracle):
    """Config 3 in miniature: L = 0.5 sum w m^2, gradients by autograd vs central FD (FP64)."""
    from ppm_legacy_amd import autograd
    N = 16
    xp = wl.jittered_lattice(N, seed=3)
    up = wl.strengths(3, len(xp), 1)
    w = np.random.default_rng(3).normal(size=(N, N, N, 1))
    mesh = core.Mesh(3, N + 1)
    x = _dev(xp, torch.float64).requires_grad_(True)
    u = _dev(up, torch.float64).requires_grad_(True)
    wt = _dev(w, torch.float64)
    m = autograd.p2m(x, u, mesh)
    loss = 0.5 * (wt * m * m).sum()
    loss.backward()
    t = oracle.topo(3, [N + 1] * 3)
    _, f = oracle.p2m(t, MP4, xp, up, [2, 2, 2])
    mref = oracle.unique_nodes(f, [2, 2, 2], 3)
    _, gu_ref, gx_ref = oracle.p2m_adjoint(t, MP4, xp, up, w * mref)
    assert _rel(u.grad.cpu().numpy(), gu_ref) < 1e-12
    assert _rel(x.grad.cpu().numpy(), gx_ref) < 1e-12
    h = 1.0 / N
    step = 1e-6 * h
    rng = np.random.default_rng(0)
    for p in rng.choice(len(xp), size=16, replace=False):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                xs = xp.copy()
                xs[p, d] += sgn * step
                _, fs = oracle.p2m(t, MP4, xs, up, [2, 2, 2])
                ms = oracle.unique_nodes(fs, [2, 2, 2], 3)
                vals.append(0.5 * (w * ms * ms).sum())
            fd = (vals[0] - vals[1]) / (2 * step)
            assert abs(fd - x.grad[p, d].item()) <= 1e-6 * max(1.0, abs(fd))


# --------------------------------------------------------------- rank3d --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", [F64, F32])
def test_util_rank_bit_exact(core, oracle, dim, prec):
    rng = np.random.default_rng(50 + dim)
    dt = np.float32 if prec == F32 else np.float64
    xp = rng.uniform(-0.1, 1.1, size=(20000, dim)).astype(dt)
    xp[::5] = xp[1::5][: len(xp[::5])]
    xp[3] = 1.0                                          # exactly on xmax
    Nc = [13, 7, 9][:dim]
    Ngl = [1, 0, 2, 1, 2, 0][: 2 * dim] if dim == 3 else [1, 0, 1, 2]
    lp_ref, lh_ref, ic_ref, nout, _ = oracle.rank(xp, [0.0] * dim, [1.0] * dim, Nc, Ngl, dtype=dt)
    lp, lh = core.util_rank(_dev(xp, torch.float32 if prec == F32 else torch.float64), [0.0] * dim, [1.0] * dim,
                            Nc, Ngl)
    assert np.array_equal(lp.cpu().numpy(), lp_ref)
    assert np.array_equal(lh.cpu().numpy(), lh_ref)
    assert nout > 0


# --------------------------------------------------------------- remesh --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
def test_comp_weights_and_remesh(core, oracle, dim, kernel):
    N = 12
    xp, up = _case(dim, N, 1000, 400 + dim + kernel, 3)
    g = [2] * dim
    t = oracle.topo(dim, [N + 1] * dim)
    info, wref = oracle.comp_weights(t, kernel, xp)
    assert info == 0
    mesh = core.Mesh(dim, N + 1)
    x = _dev(xp, torch.float64)
    w = core.comp_weights(mesh, x, kernel)
    for d in range(dim):
        assert np.max(np.abs(w[d].cpu().numpy() - wref[d])) < 1e-15
    info, fref = oracle.remesh(t, kernel, xp, up, g, wref)
    assert info == 0
    field = torch.zeros(mesh.padded_shape(3, g), dtype=torch.float64, device="cuda")
    core.remesh(mesh, x, _dev(up, torch.float64), w, field, [-2] * dim, kernel)
    assert _rel(field.cpu().numpy(), fref) < 1e-12


@pytest.mark.parametrize("order", ["lattice", "shuffled"])
def test_m2p_prebin_on_a_second_stream(core, order):
    """ppmg_m2p_prebin bins the particles of the next M2P on a side stream
    while a P2M runs on the main one; the M2P that follows (ordered after
    both) must equal an M2P that bins itself, bitwise.  A prebinning is used
    only by the M2P of the same positions: another position array, or another
    binning on the plan in between, makes the M2P bin again."""
    N = 40
    xq = wl.jittered_lattice(N, seed=31)
    if order == "shuffled":
        xq = xq[np.random.default_rng(5).permutation(len(xq))]
    xs = wl.jittered_lattice(N, seed=32)
    us = wl.strengths(33, len(xs), 3)
    mesh = core.Mesh(3, N + 1)
    x, u, q = _dev(xs, torch.float64), _dev(us, torch.float64), _dev(xq, torch.float64)
    field = torch.empty(mesh.compact_shape(3), dtype=torch.float64, device="cuda")
    core.p2m(mesh, x, u, field, [0] * 3, MP4, F64)
    ref = torch.empty((len(xq), 3), dtype=torch.float64, device="cuda")
    core.m2p(mesh, q, field, [0] * 3, ref, MP4, F64)
    torch.cuda.synchronize()
    main, side = torch.cuda.current_stream(), torch.cuda.Stream()
    plan = core.Plan()
    pp = core.Plan()
    field2 = torch.empty_like(field)
    for rep in range(3):
        out = torch.full_like(ref, 7.0)
        main.wait_stream(side)
        core.p2m(mesh, x, u, field2, [0] * 3, MP4, F64, pp)
        with torch.cuda.stream(side):
            core.m2p_prebin(mesh, q, plan, MP4, F64)
            side.wait_stream(main)
            core.m2p(mesh, q, field2, [0] * 3, out, MP4, F64, plan)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    # prebinned for another array of the same particles: the M2P bins again
    q2 = q.clone()
    core.m2p_prebin(mesh, q, plan, MP4, F64)
    out = torch.empty_like(ref)
    core.m2p(mesh, q2, field, [0] * 3, out, MP4, F64, plan)
    assert torch.equal(out, ref)
    # a P2M binning on the plan discards the prebinning
    core.m2p_prebin(mesh, q, plan, MP4, F64)
    core.p2m(mesh, x, u, field2, [0] * 3, MP4, F64, plan)
    core.m2p(mesh, q, field, [0] * 3, out, MP4, F64, plan)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


@pytest.mark.parametrize("prec", [F64, F32C])
@pytest.mark.parametrize("layout", ["padded", "compact"])
def test_remesh_tiled_deposits_the_weight_tables(core, oracle, monkeypatch, prec, layout):
    """3D M'4 ppm_rmsh_remesh (ppm_rmsh_remesh.f:618-802) runs on the tiled
    deposit with the caller's weight tables.  The tables here are M'4's
    perturbed by up to 20% per entry, so a kernel that recomputed weights
    from the positions would fail.  Clustered particles put hundreds in some
    cells (the tile kernel's later rounds and its one-at-a-time leftovers).
    Checked against the oracle's remesh with the same tables on every stored
    node (reference box with ghosts, or the compact unique nodes) and, in
    FP64, against the general path."""
    N = 20
    rng = np.random.default_rng(88 + prec)
    xp = np.concatenate([rng.normal(0.5, 0.03, size=(2500, 3)) % 1.0, rng.uniform(0, 1, size=(1500, 3))])
    up = rng.normal(size=(len(xp), 3))
    t = oracle.topo(3, [N + 1] * 3)
    info, wref = oracle.comp_weights(t, MP4, xp)
    assert info == 0
    wt = [w * (1.0 + 0.2 * rng.uniform(-1, 1, size=w.shape)) for w in wref]
    g = [2, 2, 2]
    info, fref = oracle.remesh(t, MP4, xp, up, g, wt)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    vdt = torch.float64 if prec == F64 else torch.float32
    x = _dev(xp, torch.float64)
    u = _dev(up, vdt)
    w = [_dev(a, vdt) for a in wt]
    if layout == "padded":
        shape, lo, ref = mesh.padded_shape(3, g), [-2] * 3, fref
    else:
        shape, lo, ref = mesh.compact_shape(3), [0] * 3, oracle.unique_nodes(fref, g, 3)
    field = torch.empty(shape, dtype=vdt, device="cuda")
    core.remesh(mesh, x, u, w, field, lo, MP4, prec)
    tol = 1e-12 if prec == F64 else 1e-6
    assert _rel(field.double().cpu().numpy(), ref) < tol
    if prec == F64:
        monkeypatch.setenv("PPMG_GENERAL", "1")
        gen = torch.empty_like(field)
        core.remesh(mesh, x, u, w, gen, lo, MP4, prec)
        monkeypatch.delenv("PPMG_GENERAL")
        assert _rel(field.cpu().numpy(), gen.cpu().numpy()) < 1e-12


def test_create_part(core, oracle):
    N = 10
    g = [1, 1, 1]
    t = oracle.topo(3, [N + 1] * 3)
    f = np.random.default_rng(6).normal(size=oracle.field_shape(t, 3, g))
    cut = [0.7, 1.9]
    info, xref, uref = oracle.create_part(t, f, g, cut)
    assert info == 0
    mesh = core.Mesh(3, N + 1)
    xp, up = core.create_part(mesh, _dev(f, torch.float64), [-1] * 3, cut)
    assert np.array_equal(up.cpu().numpy(), uref)
    assert np.array_equal(xp.cpu().numpy(), xref)
    cw = [0.5, 2.0, 1.0]
    info, xref, uref = oracle.create_part(t, f, g, cut, cutoff_weights=np.array(cw))
    xp, up = core.create_part(mesh, _dev(f, torch.float64), [-1] * 3, cut, cutoff_weights=cw)
    assert np.array_equal(up.cpu().numpy(), uref)


def test_create_part_slave_field(core, oracle):
    """field_wp / wp (ppm_rmsh_create_part.f:626-632): the slave values at the
    selected nodes.  The slave is an affine image of the master with its own
    ghost width and component count, so wp must equal that image of up, and
    the positions must match the node lattice of the oracle."""
    N = 9
    t = oracle.topo(3, [N + 1] * 3)
    f = np.random.default_rng(16).normal(size=oracle.field_shape(t, 3, [1, 1, 1]))
    cut = [0.5, 1.5]
    info, xref, uref = oracle.create_part(t, f, [1, 1, 1], cut)
    assert info == 0 and len(uref) > 0
    mesh = core.Mesh(3, N + 1)
    # slave: 2 components, ghost 2, values (2*u0 + 1, -u2) at every stored node
    inner = f[1:-1, 1:-1, 1:-1]
    s = np.zeros((N + 5, N + 5, N + 5, 2))
    s[2:-2, 2:-2, 2:-2, 0] = 2.0 * inner[..., 0] + 1.0
    s[2:-2, 2:-2, 2:-2, 1] = -inner[..., 2]
    xp, up, wp = core.create_part(mesh, _dev(f, torch.float64), [-1] * 3, cut, slave=_dev(s, torch.float64),
                                  slave_lo=[-2] * 3)
    assert np.array_equal(xp.cpu().numpy(), xref)
    assert np.array_equal(up.cpu().numpy(), uref)
    w = wp.cpu().numpy()
    assert np.array_equal(w[:, 0], 2.0 * uref[:, 0] + 1.0)
    assert np.array_equal(w[:, 1], -uref[:, 2])
    # without resetpos there are no positions, as the reference leaves xp alone
    xn, un, wn = core.create_part(mesh, _dev(f, torch.float64), [-1] * 3, cut, resetpos=False,
                                  slave=_dev(s, torch.float64), slave_lo=[-2] * 3)
    assert xn is None and np.array_equal(un.cpu().numpy(), uref) and np.array_equal(wn.cpu().numpy(), w)


@pytest.mark.parametrize("bc", [[1, 1, 1, 1, 2, 2], [2, 2, 1, 1, 1, 1], [2] * 6], ids=["open_z", "open_x", "all_open"])
def test_remesh_create_part_bc_on_open_axes(core, oracle, bc):
    """comp_weights + remesh, create_part and impose_part_bc on meshes with
    non-periodic axes, against the oracle: remesh folds only the periodic
    axes, create_part takes node N of an open axis (a real node there) but
    not of a periodic one, impose_part_bc moves only periodic coordinates."""
    N = 10
    rng = np.random.default_rng(70 + sum(bc))
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    mesh = core.Mesh(3, N + 1, bcdef=bc)
    xp = rng.uniform(0, 1, size=(1500, 3))
    up = rng.normal(size=(len(xp), 2))
    info, wref = oracle.comp_weights(t, MP4, xp)
    assert info == 0
    x = _dev(xp, torch.float64)
    w = core.comp_weights(mesh, x, MP4)
    for d in range(3):
        assert np.max(np.abs(w[d].cpu().numpy() - wref[d])) < 1e-15
    info, fref = oracle.remesh(t, MP4, xp, up, [2] * 3, wref)
    assert info == 0
    field = torch.zeros(mesh.padded_shape(2, [2] * 3), dtype=torch.float64, device="cuda")
    core.remesh(mesh, x, _dev(up, torch.float64), w, field, [-2] * 3, MP4)
    assert _rel(field.cpu().numpy(), fref) < 1e-12
    f = rng.normal(size=oracle.field_shape(t, 3, [1] * 3))
    info, xref, uref = oracle.create_part(t, f, [1] * 3, [0.7, 1.9])
    assert info == 0 and len(xref) > 0
    xc, uc = core.create_part(mesh, _dev(f, torch.float64), [-1] * 3, [0.7, 1.9])
    assert np.array_equal(xc.cpu().numpy(), xref)
    assert np.array_equal(uc.cpu().numpy(), uref)
    xb = rng.uniform(-0.5, 1.5, size=(2000, 3))
    ref = oracle.impose_part_bc(t, xb)
    xd = _dev(xb, torch.float64)
    core.impose_part_bc(mesh, xd)
    assert np.array_equal(xd.cpu().numpy(), ref)


@pytest.mark.parametrize("bc", [[1] * 6, [2] * 6, [1, 1, 1, 1, 2, 2], [2, 2, 1, 1, 1, 1]],
                         ids=["periodic_box", "all_open", "open_z", "open_x"])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
@pytest.mark.parametrize("prec", [F64, F32C])
def test_adjoint_stored_boxes(core, oracle, bc, kernel, prec):
    """The adjoint with the cotangent on the reference's stored box (ghosts
    and node N included), periodic, open and mixed: each particle collects g
    from every stored node the forward writes it into.  Strength gradient:
    <P2M(u), g> = <u, dL/du> over every stored node, and per particle
    dL/du_p,c = <P2M({x_p}, {e_c}), g> with the oracle's P2M.  Position
    gradient: central differences of the oracle's P2M of one particle,
    L_p(x) = <P2M({x}, {u_p}), g>, with step 1e-6 h (both kernels are C0 at
    worst only on nodes, which random positions miss).  FP32c (values in FP32,
    positions in FP64) against the same FP64 references on the FP32-rounded
    strengths and cotangent: 1e-6 for the sums, 1e-5 of the gradient scale
    for the differences."""
    N = 10
    rng = np.random.default_rng(80 + kernel + sum(bc))
    xp = rng.uniform(0.05, 0.95, size=(600, 3))
    up = rng.normal(size=(len(xp), 2))
    gc = rng.normal(size=(N + 5,) * 3 + (2,))
    if prec == F32C:
        up = up.astype(np.float32).astype(np.float64)
        gc = gc.astype(np.float32).astype(np.float64)
    tol_sum, tol_fd = (1e-12, 1e-6) if prec == F64 else (1e-6, 1e-5)
    vdt = _val_dtype(prec)
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    mesh = core.Mesh(3, N + 1, bcdef=bc)
    x, u = _dev(xp, torch.float64), _dev(up, vdt)
    field = torch.empty(mesh.padded_shape(2, [2] * 3), dtype=vdt, device="cuda")
    plan = core.p2m(mesh, x, u, field, [-2] * 3, kernel, prec)
    gu, gx = core.p2m_adjoint(mesh, x, u, _dev(gc, vdt), [-2] * 3, plan, kernel, prec)
    gu, gx = gu.cpu().numpy().astype(np.float64), gx.cpu().numpy().astype(np.float64)
    fnp = field.cpu().numpy().astype(np.float64)
    lhs = float(np.sum(fnp * gc))
    rhs = float(np.sum(up * gu))
    assert abs(lhs - rhs) <= tol_sum * np.sum(np.abs(fnp * gc))
    for p in range(0, len(xp), 50):
        for c in range(2):
            e = np.zeros((1, 2))
            e[0, c] = 1.0
            info, fp = oracle.p2m(t, kernel, xp[p:p + 1], e, [2] * 3)
            assert info == 0
            assert abs(gu[p, c] - np.sum(fp * gc)) <= tol_sum * np.max(np.abs(gu)), (p, c)
    h = 1e-6 / N
    scale = np.max(np.abs(gx))
    for p in range(0, len(xp), 25):
        for d in range(3):
            xs = []
            for s in (1.0, -1.0):
                xq = xp[p:p + 1].copy()
                xq[0, d] += s * h
                info, fp = oracle.p2m(t, kernel, xq, up[p:p + 1], [2] * 3)
                assert info == 0
                xs.append(float(np.sum(fp * gc)))
            fd = (xs[0] - xs[1]) / (2 * h)
            assert abs(gx[p, d] - fd) <= tol_fd * scale, (p, d, gx[p, d], fd)


@pytest.mark.parametrize("prec", [F64, F32C])
@pytest.mark.parametrize("layout", ["box_g2", "mixed", "compact"])
def test_adjoint_tiled_stored_box_vs_general(core, monkeypatch, prec, layout):
    """The tiled adjoint reads a cotangent on a stored box of periodic axes
    (the reference's 1-g..Nm+g box, or compact on some axes and boxed with
    other ghost widths on others) through the copies of each unique node
    and the raw ghost nodes.  Clustered particles (crowded cells, particles
    on the faces' stencils) with lda 3: strength and position gradients
    against the general per-particle adjoint, which collects g node by node
    from every stored node the forward writes (1e-12 in FP64, 1e-6 in
    FP32c)."""
    N = 24
    rng = np.random.default_rng(140 + prec)
    xp = np.concatenate([rng.normal(0.02, 0.03, size=(2000, 3)) % 1.0, rng.uniform(0, 1, size=(2000, 3))])
    up = rng.normal(size=(len(xp), 3))
    vdt = _val_dtype(prec)
    mesh = core.Mesh(3, N + 1)
    if layout == "box_g2":
        lo, ext = [-2, -2, -2], [N + 5] * 3
    elif layout == "mixed":
        lo, ext = [0, -1, -3], [N, N + 3, N + 7]
    else:
        lo, ext = [0, 0, 0], [N] * 3
    x, u = _dev(xp, torch.float64), _dev(up, vdt)
    field = torch.empty((ext[2], ext[1], ext[0], 3), dtype=vdt, device="cuda")
    plan = core.p2m(mesh, x, u, field, lo, MP4, prec)
    g = _dev(rng.normal(size=(ext[2], ext[1], ext[0], 3)), vdt)
    gu, gx = core.p2m_adjoint(mesh, x, u, g, lo, plan, MP4, prec)
    monkeypatch.setenv("PPMG_GENERAL", "1")
    ru, rx = core.p2m_adjoint(mesh, x, u, g, lo, plan, MP4, prec)
    monkeypatch.delenv("PPMG_GENERAL")
    tol = 1e-12 if prec == F64 else 1e-6
    assert _rel(gu.double().cpu().numpy(), ru.double().cpu().numpy()) < tol
    assert _rel(gx.double().cpu().numpy(), rx.double().cpu().numpy()) < tol


def test_impose_part_bc(core, oracle):
    t = oracle.topo(3, [9, 9, 9], [-1.0, 0.0, 0.0], [1.0, 1.0, 2.0])
    rng = np.random.default_rng(12)
    xp = rng.uniform(-1.5, 2.5, size=(5000, 3))
    ref = oracle.impose_part_bc(t, xp)
    mesh = core.Mesh(3, 9, [-1.0, 0.0, 0.0], [1.0, 1.0, 2.0])
    x = _dev(xp, torch.float64)
    core.impose_part_bc(mesh, x)
    assert np.array_equal(x.cpu().numpy(), ref)


# ------------------------------------------------------------ generators --
def test_device_generators_match_host(core):
    N = 24
    xd = core.gen_jittered_lattice(N, seed=77).cpu().numpy()
    assert np.array_equal(xd, wl.jittered_lattice(N, seed=77))
    ud = core.gen_normals(1000, 3, 77).cpu().numpy()
    assert np.allclose(ud, wl.strengths(77, 1000, 3), rtol=1e-13, atol=1e-13)
    modes = wl.fourier_modes(9)
    vd = core.gen_divfree_velocity(core.gen_jittered_lattice(8, 1), modes).cpu().numpy()
    assert np.allclose(vd, wl.divfree_velocity(wl.jittered_lattice(8, 1), modes), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------- registry --
def test_registry_host_arrays_like_the_reference(oracle):
    """ppm_init / ppm_mktopo / ppm_interp_p2m / m2p with host (numpy) arrays."""
    from ppm_legacy_amd import ppm
    assert ppm.ppm_init(3, ppm.ppm_kind_double, -15, None, 0) == 0
    N = 16
    g = [2, 2, 2]
    info, topo, meshid = ppm.ppm_mktopo([0.0] * 3, [1.0] * 3, [1] * 6, [N + 1] * 3, g)
    assert info == 0
    xp = wl.jittered_lattice(N, seed=21)
    up = wl.strengths(21, len(xp), 3)
    field = np.zeros((N + 1 + 4,) * 3 + (3,))
    assert ppm.ppm_interp_p2m(xp, len(xp), up, topo, meshid, ppm.ppm_param_rmsh_kernel_mp4, g, field) == 0
    t = oracle.topo(3, [N + 1] * 3)
    _, ref = oracle.p2m(t, MP4, xp, up, g)
    assert _rel(field, ref) < 1e-12
    back = np.zeros_like(up)
    assert ppm.ppm_interp_m2p(xp, len(xp), back, topo, meshid, ppm.ppm_param_rmsh_kernel_mp4, g, field) == 0
    _, bref = oracle.m2p(t, MP4, xp, ref, g)
    assert _rel(back, bref) < 1e-12
    # single specifics
    xs, us = xp.astype(np.float32), up.astype(np.float32)
    fs = np.zeros(field.shape, np.float32)
    assert ppm.ppm_interp_p2m(xs, len(xs), us, topo, meshid, ppm.ppm_param_rmsh_kernel_mp4, g, fs) == 0
    _, sref = oracle.p2m(t, MP4, xs, us, g, dtype=np.float32)
    assert _rel(fs, sref) < 1e-6
    # FP32-compensated form behind the unchanged double signature
    assert ppm.ppm_set_precision_mode(ppm.PREC_F32C) == 0
    fc = np.zeros_like(field)
    assert ppm.ppm_interp_p2m(xp, len(xp), up, topo, meshid, ppm.ppm_param_rmsh_kernel_mp4, g, fc) == 0
    assert _rel(fc, ref) < 1e-6
    assert ppm.ppm_set_precision_mode(ppm.PREC_F64) == 0
    # the reference's fatal error on a particle outside the sub
    bad = xp.copy()
    bad[5, 1] = 1.0
    assert ppm.ppm_interp_p2m(bad, len(bad), up, topo, meshid, 2, g, field) == -1
    assert ppm.ppm_finalize() == 0


@pytest.mark.parametrize("bc", [[1, 1, 1, 1, 2, 2], [2, 2, 2, 2, 1, 1]], ids=["open_z", "open_xy"])
@pytest.mark.parametrize("kernel", [MP4, BSP2])
def test_registry_mixed_axes(oracle, bc, kernel):
    """The registry entry points (host arrays, the Fortran drop-in's route) on
    meshes with open and periodic axes: P2M, and M2P of a random stored field
    whose ghosts all differ, against the oracle, in the double and single
    specifics.  Particles sit on the open faces too."""
    from ppm_legacy_amd import ppm
    assert ppm.ppm_init(3, ppm.ppm_kind_double, -15, None, 0) == 0
    N = 14
    g = [2, 2, 2]
    info, topo, meshid = ppm.ppm_mktopo([0.0] * 3, [1.0] * 3, bc, [N + 1] * 3, g)
    assert info == 0
    rng = np.random.default_rng(90 + kernel + sum(bc))
    xp = rng.uniform(0, 1, size=(5000, 3))
    for d in range(3):
        if bc[2 * d] != 1:
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 1.0
            xp[rng.uniform(size=len(xp)) < 0.05, d] = 0.0
    up = rng.normal(size=(len(xp), 2))
    t = oracle.topo(3, [N + 1] * 3, bcdef=bc)
    shape = (N + 1 + 4,) * 3 + (2,)
    fr = rng.normal(size=shape)
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        x, u = xp.astype(dt), up.astype(dt)
        field = np.zeros(shape, dt)
        assert ppm.ppm_interp_p2m(x, len(x), u, topo, meshid, kernel, g, field) == 0
        info, ref = oracle.p2m(t, kernel, x, u, g, dtype=dt)
        assert info == 0
        assert _rel(field, ref) < tol
        back = np.zeros_like(u)
        assert ppm.ppm_interp_m2p(x, len(x), back, topo, meshid, kernel, g, fr.astype(dt).copy()) == 0
        info, bref = oracle.m2p(t, kernel, x, fr.astype(dt), g, dtype=dt)
        assert info == 0
        assert _rel(back, bref) < tol
    assert ppm.ppm_finalize() == 0


# ---------------------------------------------------------- ghost mapping --
@pytest.mark.parametrize("dim,bc", [(3, [1] * 6), (3, [1, 1, 2, 2, 1, 1]), (2, [1] * 4), (2, [2, 2, 1, 1])])
@pytest.mark.parametrize("g", [0, 1, 3])
@pytest.mark.parametrize("prec,lda", [(F64, 3), (F64, 1), (F32, 2)])
def test_map_field_ghost_vs_oracle(core, oracle, dim, bc, g, prec, lda):
    """ppm_map_field_ghost ghost_put and ghost_get (push, send, pop) on a
    padded field of random values, periodic and non-periodic axes, against
    the oracle's restatement of the reference's block lists: bit-exact (the
    blocks are applied in the same order, so every node sees the same
    additions in the same order)."""
    from ppm_legacy_amd import _lib as L
    nm = [9, 7, 6][:dim]
    rng = np.random.default_rng(5 + g + 10 * dim + lda)
    t = oracle.topo(dim, nm, bcdef=bc)
    mesh = core.Mesh(dim, nm, bcdef=bc)
    shape = tuple(reversed([n + 2 * g for n in nm])) + (lda,)
    vals = rng.normal(size=shape).astype(_np_val(prec) if prec != F32 else np.float32)
    f = _dev(vals, _val_dtype(prec) if prec != F32 else torch.float32)
    core.map_field_ghost(mesh, f, [-g] * dim, [g] * dim, L.MAP_GHOST_PUT, prec)
    info, ref = oracle.ghost_put(t, vals, [g] * dim)
    assert info == 0
    assert np.array_equal(f.cpu().numpy(), ref)
    core.map_field_ghost(mesh, f, [-g] * dim, [g] * dim, L.MAP_GHOST_GET, prec)
    info, ref = oracle.ghost_get(t, ref, [g] * dim)
    assert info == 0
    assert np.array_equal(f.cpu().numpy(), ref)


def test_map_field_ghost_get_then_m2p_equals_m2p(core, oracle):
    """M2P reads periodic images through ghost_get's node map, so an explicit
    ghost_get on the padded field changes nothing it reads; after ghost_get
    every periodic image node equals its real node."""
    from ppm_legacy_amd import _lib as L
    N, g = 12, 2
    rng = np.random.default_rng(3)
    xp = rng.uniform(0, 1, size=(2000, 3))
    up = rng.normal(size=(2000, 2))
    mesh = core.Mesh(3, N + 1)
    f = torch.zeros(mesh.padded_shape(2, [g] * 3), dtype=torch.float64, device="cuda")
    core.p2m(mesh, _dev(xp, torch.float64), _dev(up, torch.float64), f, [-g] * 3)
    before = torch.empty((2000, 2), dtype=torch.float64, device="cuda")
    core.m2p(mesh, _dev(xp, torch.float64), f, [-g] * 3, before)
    core.map_field_ghost(mesh, f, [-g] * 3, [g] * 3, L.MAP_GHOST_GET)
    after = torch.empty_like(before)
    core.m2p(mesh, _dev(xp, torch.float64), f, [-g] * 3, after)
    assert torch.equal(before, after)
    h = f.cpu().numpy()
    # node i and i + N are images of each other on every axis (box index = node + g)
    assert np.array_equal(h[:, :, g:g + 3], h[:, :, g + N:g + N + 3])
    assert np.array_equal(h[:, g:g + 3], h[:, g + N:g + N + 3])
    assert np.array_equal(h[g:g + 3], h[g + N:g + N + 3])


def test_map_field_ghost_rejects_bad_arguments(core):
    from ppm_legacy_amd import _lib as L
    mesh = core.Mesh(3, 9)
    f = torch.zeros(mesh.padded_shape(1, [1] * 3), dtype=torch.float64, device="cuda")
    with pytest.raises(L.PpmgError) as e:
        core.map_field_ghost(mesh, f, [-1] * 3, [2] * 3, L.MAP_GHOST_GET)   # box holds only one ghost layer
    assert e.value.info == -2 and e.value.code == L.ERR_ARGUMENT
    with pytest.raises(L.PpmgError) as e:
        core.map_field_ghost(mesh, f, [-1] * 3, [1] * 3, 7)
    assert e.value.code == L.ERR_ARGUMENT


def _kind_values(rng, kind, shape, L):
    """Random values of an element kind: (numpy array, torch dtype)."""
    if kind == L.KIND_I:
        return rng.integers(-2 ** 31, 2 ** 31 - 1, size=shape, dtype=np.int64).astype(np.int32), torch.int32
    if kind == L.KIND_L:
        return (rng.random(size=shape) < 0.7).astype(np.int32) * rng.integers(1, 4, size=shape, dtype=np.int32), \
            torch.int32
    if kind == L.KIND_DC:
        return rng.normal(size=shape) + 1j * rng.normal(size=shape), torch.complex128
    if kind == L.KIND_SC:
        return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64), torch.complex64
    if kind == L.KIND_S:
        return rng.normal(size=shape).astype(np.float32), torch.float32
    return rng.normal(size=shape), torch.float64


@pytest.mark.parametrize("dim,bc,g", [(3, [1] * 6, 2), (3, [1, 1, 2, 2, 1, 1], 1), (2, [1] * 4, 3),
                                      (2, [2, 2, 1, 1], 2)])
@pytest.mark.parametrize("kind", [0, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("masked", [False, True])
def test_map_field_ghost_kinds_vs_oracle(core, oracle, dim, bc, g, kind, masked):
    """ppm_map_field_ghost for every element kind of the reference's specifics
    (double, single, integer, logical, double and single complex), with and
    without the mask, ghost_put then ghost_get, against the oracle's
    restatement of the reference's push into a double buffer and per-kind pop
    (ppm_map_field_pop_3d.f:1818-1897): bit-exact, integers wrapping on
    overflow, logicals combined by .AND."""
    from ppm_legacy_amd import _lib as L
    nm = [9, 7, 6][:dim]
    rng = np.random.default_rng(41 + 7 * kind + g + 10 * dim + masked)
    t = oracle.topo(dim, nm, bcdef=bc)
    mesh = core.Mesh(dim, nm, bcdef=bc)
    shape = tuple(reversed([n + 2 * g for n in nm])) + (2,)
    vals, tdt = _kind_values(rng, kind, shape, L)
    mask = (rng.random(size=shape[:-1]) < 0.6).astype(np.int32) if masked else None
    f = torch.as_tensor(np.ascontiguousarray(vals)).to(tdt).cuda()
    m = torch.as_tensor(mask).cuda() if masked else None
    core.map_field_ghost_kind(mesh, f, [-g] * dim, [g] * dim, L.MAP_GHOST_PUT, kind, m)
    info, ref = oracle.ghost_kind(t, vals, [g] * dim, kind, False, mask)
    assert info == 0
    assert np.array_equal(f.cpu().numpy(), ref)
    core.map_field_ghost_kind(mesh, f, [-g] * dim, [g] * dim, L.MAP_GHOST_GET, kind, m)
    info, ref = oracle.ghost_kind(t, ref, [g] * dim, kind, True, mask)
    assert info == 0
    assert np.array_equal(f.cpu().numpy(), ref)


@pytest.mark.parametrize("on_device", [False, True])
def test_ppm_map_field_ghost_sequence_kinds(oracle, on_device):
    """The mirror's ppm_map_field_ghost call sequence (init, ghost_get or
    ghost_put, push, send, pop) on host arrays and on device tensors of every
    kind: a LOGICAL field as a bool array, a complex one as complex, with a
    bool mask; results as the oracle's.  Out of sequence calls and a field
    type with no specific are errors."""
    from ppm_legacy_amd import _lib as L
    from ppm_legacy_amd import ppm
    N, g = 8, 2
    assert ppm.ppm_init(3, ppm.ppm_kind_double, -15, None, 0) == 0
    try:
        info, topo, meshid = ppm.ppm_mktopo([0.0] * 3, [1.0] * 3, [1, 1, 2, 2, 1, 1], [N + 1] * 3, [g] * 3)
        assert info == 0
        t = oracle.topo(3, [N + 1] * 3, bcdef=[1, 1, 2, 2, 1, 1])
        rng = np.random.default_rng(9)
        shape = (N + 1 + 2 * g,) * 3 + (3,)
        mask = rng.random(size=shape[:-1]) < 0.5
        cases = [(L.KIND_D, rng.normal(size=shape)),
                 (L.KIND_I, rng.integers(-1000, 1000, size=shape).astype(np.int32)),
                 (L.KIND_L, rng.random(size=shape) < 0.8),
                 (L.KIND_DC, rng.normal(size=shape) + 1j * rng.normal(size=shape))]
        wrap = (lambda a: torch.as_tensor(a).cuda()) if on_device else (lambda a: a.copy())
        back = (lambda a: a.cpu().numpy()) if on_device else (lambda a: a)
        for kind, vals in cases:
            for mt in (ppm.ppm_param_map_ghost_put, ppm.ppm_param_map_ghost_get):
                for use_mask in (False, True):
                    fv = wrap(vals)
                    mk = wrap(mask) if use_mask else None
                    for step in (ppm.ppm_param_map_init, mt, ppm.ppm_param_map_push, ppm.ppm_param_map_send):
                        assert ppm.ppm_map_field_ghost(fv, topo, meshid, [g] * 3, step, lda=3) == 0
                    assert ppm.ppm_map_field_ghost(fv, topo, meshid, [g] * 3, ppm.ppm_param_map_pop, lda=3,
                                                   mask=mk) == 0
                    ref_in = vals.astype(np.int32) if kind == L.KIND_L else vals
                    info, ref = oracle.ghost_kind(t, ref_in, [g] * 3, kind, mt == ppm.ppm_param_map_ghost_get,
                                                  mask.astype(np.int32) if use_mask else None)
                    assert info == 0
                    got = back(fv)
                    assert np.array_equal(got.astype(np.int32) if kind == L.KIND_L else got, ref), (kind, mt, use_mask)
        # pop without push; a field type with no specific
        fv = wrap(cases[0][1])
        assert ppm.ppm_map_field_ghost(fv, topo, meshid, [g] * 3, ppm.ppm_param_map_pop, lda=3) != 0
        for step in (ppm.ppm_param_map_ghost_get, ppm.ppm_param_map_push):
            assert ppm.ppm_map_field_ghost(fv, topo, meshid, [g] * 3, step, lda=3) == 0
        bad = wrap(np.zeros(shape, dtype=np.int16))
        assert ppm.ppm_map_field_ghost(bad, topo, meshid, [g] * 3, ppm.ppm_param_map_pop, lda=3) != 0
    finally:
        assert ppm.ppm_finalize() == 0


def test_map_field_ghost_kind_rejects_bad_arguments(core):
    from ppm_legacy_amd import _lib as L
    mesh = core.Mesh(3, 9)
    f = torch.zeros(mesh.padded_shape(1, [1] * 3), dtype=torch.int32, device="cuda")
    with pytest.raises(L.PpmgError) as e:
        core.map_field_ghost_kind(mesh, f, [-1] * 3, [2] * 3, L.MAP_GHOST_PUT, L.KIND_I)   # one ghost layer held
    assert e.value.code == L.ERR_ARGUMENT
    with pytest.raises(L.PpmgError) as e:
        plan = core.Plan()
        fd = core.field_desc(mesh, f, [-1] * 3)
        g = (ctypes.c_int * 3)(1, 1, 1)
        L.check(L.lib().ppmg_map_field_ghost_kind(mesh.ref(), 9, ctypes.byref(fd), g, L.MAP_GHOST_PUT, None,
                                                  plan.handle, None))
    assert e.value.code == L.ERR_ARGUMENT
    with pytest.raises(ValueError):
        core.map_field_ghost_kind(mesh, f, [-1] * 3, [1] * 3, L.MAP_GHOST_PUT, L.KIND_I,
                                  torch.zeros((3, 3, 3), dtype=torch.int32, device="cuda"))


# --------------------------------------------------------------- util_sort --
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_util_sort_vs_rank(core, oracle, dim, dtype):
    """ppm_util_sort3d / sort2d: the rows of xp in the stable order of
    ppm_util_rank with no ghost layers (the oracle's ranking), npbx the box
    counts; both through the core API and the reference-named host mirror."""
    from ppm_legacy_amd import ppm
    rng = np.random.default_rng(dim)
    ncells = [7, 5, 4][:dim]
    xmin, xmax = [-1.0] * dim, [2.0] * dim
    xp = rng.uniform(-1.0, 2.0, size=(5000, dim)).astype(np.float32 if dtype == torch.float32 else np.float64)
    odt = np.float32 if dtype == torch.float32 else np.float64
    lref, href, ic, _, _ = oracle.rank(xp, xmin, xmax, ncells, [0] * (2 * dim), dtype=odt)
    assert ic == len(xp)
    x = _dev(xp, dtype)
    npbx = core.util_sort(x, xmin, xmax, ncells)
    assert np.array_equal(x.cpu().numpy(), xp[lref - 1])
    assert np.array_equal(npbx.cpu().numpy(), np.diff(href))
    # host arrays through the registry
    xh = xp.copy()
    nh = np.zeros(int(np.prod(ncells)), np.int32)
    assert ppm.ppm_util_sort3d(xh, len(xh), xmin, xmax, ncells, nh) == 0
    assert np.array_equal(xh, xp[lref - 1]) and np.array_equal(nh, np.diff(href))


def test_util_sort_particle_outside_is_part_lost(core):
    from ppm_legacy_amd import _lib as L
    xp = np.random.default_rng(0).uniform(0, 1, size=(100, 3))
    xp[17] = [0.5, 1.5, 0.5]
    with pytest.raises(L.PpmgError) as e:
        core.util_sort(_dev(xp, torch.float64), [0.0] * 3, [1.0] * 3, [4, 4, 4])
    assert e.value.info == -2 and e.value.code == L.ERR_PART_LOST
