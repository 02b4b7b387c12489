"""Pins the CPU oracle: reference fixtures, known answers and invariants.

The fixtures under tests/golden/ were produced by the reference's own code
(tests/golden/make_golden.py); the oracle must reproduce them bit for bit.
"""

import math
import os
import tempfile

import numpy as np
import pytest

from oplus2_matlab_amd import mesh as M
from oracle import oracle as O


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


@pytest.fixture(scope="module")
def small_mesh():
    return M.o_mesh(24, 12, "random", seed=1)


# ---------------------------------------------------------------------------
# reference fixtures, bit for bit

@pytest.mark.parametrize("form", ["f64", "f32"])
def test_airfoil_seq_matches_reference(golden_dir, small_mesh, form):
    fx = _load(golden_dir, "airfoil_24x12_random_s1.npz")
    niter = int(fx["niter"])
    st, hist = O.airfoil_seq(small_mesh, fx[f"{form}_consts"], niter, form)
    assert np.array_equal(hist, fx[f"{form}_seq_rms"])
    for k in ("q", "qold", "adt", "res"):
        assert np.array_equal(getattr(st, k).astype(np.float64), fx[f"{form}_seq_{k}"]), k


@pytest.mark.parametrize("form", ["f64", "f32"])
def test_airfoil_omp_matches_reference(golden_dir, small_mesh, form):
    fx = _load(golden_dir, "airfoil_24x12_random_s1.npz")
    niter = int(fx["niter"])
    st, hist, _ = O.airfoil_omp(small_mesh, fx[f"{form}_consts"], niter, form,
                                part_size=64, nthreads=4)
    assert np.array_equal(hist, fx[f"{form}_omp_p64_t4_rms"])
    assert np.array_equal(st.q.astype(np.float64), fx[f"{form}_omp_p64_t4_q"])
    # the schedules differ only in the association of increments and of the
    # rms partial sums
    rel = np.abs(hist - fx[f"{form}_seq_rms"]) / np.abs(fx[f"{form}_seq_rms"])
    assert rel.max() < (1e-13 if form == "f64" else 1e-5)


@pytest.mark.slow
def test_airfoil_180k_1000_iterations_matches_reference(golden_dir):
    fx = _load(golden_dir, "airfoil_600x300_random_s1_f64.npz")
    mesh = M.o_mesh(600, 300, "random", seed=1)
    st, hist = O.airfoil_seq(mesh, fx["consts"], int(fx["niter"]), "f64")
    assert np.array_equal(hist, fx["rms"])
    s = fx["sample"]
    assert np.array_equal(st.q[s], fx["q_sample"])
    assert np.array_equal(st.adt[s], fx["adt_sample"])
    assert np.allclose(st.q.sum(axis=0), fx["q_sum"], rtol=1e-14, atol=0)


def test_unchanged_airfoil_program_rms(golden_dir):
    """The shipped float program's printed rms (every 100 iterations) is
    reproduced by the f32 oracle on the same new_grid.dat."""
    lines = open(os.path.join(golden_dir, "airfoil_ref_120x60_cartesian.txt")).read().splitlines()
    printed = {}
    for ln in lines:
        parts = ln.split()
        if len(parts) == 2 and parts[0].isdigit():
            printed[int(parts[0])] = parts[1]
    assert sorted(printed) == list(range(100, 1001, 100))
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "new_grid.dat")
        M.write_grid_dat(M.o_mesh(120, 60, "cartesian"), path)
        mesh = M.read_grid_dat(path)
    c = M.flow_constants_f32_reference().as_array()
    _, hist = O.airfoil_seq(mesh, c, 1000, "f32")
    for it, txt in printed.items():
        assert "%10.5e" % np.float32(hist[it - 1]) == "%10.5e" % float(txt) or \
            ("%10.5e" % hist[it - 1]).strip() == txt, it


def test_jacobi_matches_reference(golden_dir):
    fx = _load(golden_dir, "jacobi_64_random_s1_f64.npz")
    prob = M.jacobi_grid(64, seed=1)
    assert np.array_equal(prob.pp, fx["pp"]) and np.array_equal(prob.A, fx["A"])
    n = int(fx["niter"])
    u, du, hs, hm = O.jacobi_seq(prob, n, 0.25, -0.25, "f64")
    assert np.array_equal(hs, fx["seq_u_sum"]) and np.array_equal(hm, fx["seq_u_max"])
    assert np.array_equal(u, fx["seq_u"]) and np.array_equal(du, fx["seq_du"])
    u, du, hs, hm = O.jacobi_omp(prob, n, 0.25, -0.25, "f64", part_size=1024, nthreads=4)
    assert np.array_equal(hs, fx["omp_p1024_t4_u_sum"])
    assert np.array_equal(hm, fx["omp_p1024_t4_u_max"])
    assert np.array_equal(u, fx["omp_p1024_t4_u"])


def test_jac1_known_answer(golden_dir):
    """jac1 as shipped (NN=6, two iterations) prints exact dyadic values."""
    txt = open(os.path.join(golden_dir, "jac1_ref_nn6.txt")).read()
    prob = M.jacobi_reference_problem(6)
    u, _, hs, hm = O.jacobi_seq(prob, 2, 1.0, 1.0, "f32")
    for it in range(2):
        line = " u max/rms = %f %f " % (np.float32(hm[it]),
                                         math.sqrt(np.float32(hs[it]) / prob.nnode))
        assert line in txt
    table = []
    for j in range(5, 0, -1):
        table.append("".join(" %7.4f" % u[i - 1 + (j - 1) * 5] for i in range(1, 6)))
    assert "\n".join(table) in txt
    assert hm[0] == 0.5 and hm[1] == 0.625


def test_plans_match_reference(golden_dir, small_mesh):
    fx = _load(golden_dir, "plans_24x12_random_s1.npz")
    prob = M.jacobi_grid(16, seed=1)
    for rb in (4, 8):
        for part in (0, 64):
            for name in ("adt_calc", "res_calc", "bres_calc", "jac_res"):
                pre = f"r{rb}_p{part}_{name}_"
                if name == "jac_res":
                    n, args, ninds = O.jacobi_plan_args(prob, rb)
                else:
                    n, args, ninds = O.airfoil_plan_args(small_mesh, name, rb)
                p = O.build_plan(n, args, ninds, part)
                nb, nc, nsh = fx[pre + "header"][:3]
                assert (p["nblocks"], p["ncolors"], p["nshared"]) == (nb, nc, nsh), pre
                assert np.array_equal(np.array([p["transfer"], p["transfer2"]], np.float32),
                                      fx[pre + "transfer"]), pre
                for k in ("nthrcol", "thrcol", "offset", "nelems", "ncolblk", "blkmap",
                          "ind_offs", "ind_sizes", "nindirect"):
                    assert np.array_equal(p[k], fx[pre + k]), pre + k
                for m in range(ninds):
                    assert np.array_equal(p["ind_maps"][m], fx[pre + f"ind_map{m}"])
                for a, lm in enumerate(p["loc_maps"]):
                    if lm is not None:
                        assert np.array_equal(lm, fx[pre + f"loc_map{a}"])
                assert p["check_errors"] == 0


# ---------------------------------------------------------------------------
# invariants the mathematics fixes

def _free_stream_residual(mesh, consts, form="f64"):
    real = np.float64 if form == "f64" else np.float32
    st = O.AirfoilState(mesh, consts, form)
    O.adt_calc(form, mesh, st.x, st.q, st.adt, consts)
    O.res_calc(form, mesh, st.x, st.q, st.adt, st.res, consts)
    del real
    return st


def test_free_stream_preserved_on_interior_cells():
    """With q = qinf everywhere, the interior flux sums to zero around every
    cell that has no boundary edge (closed polygon, no dissipation)."""
    mesh = M.o_mesh(48, 24, "random", seed=3)
    c = M.flow_constants().as_array()
    st = _free_stream_residual(mesh, c)
    interior = np.ones(mesh.ncell, bool)
    interior[mesh.becell] = False
    scale = np.abs(c[6:]).max() * np.abs(mesh.x).max() * 1e-13
    assert np.abs(st.res[interior]).max() < scale
    assert np.abs(st.res[~interior]).max() > 1e3 * scale  # boundary cells do see flux


def test_free_stream_detects_flipped_edge():
    """A single mis-oriented edge breaks free-stream preservation (the check
    above has teeth)."""
    mesh = M.o_mesh(48, 24, "random", seed=3)
    k = int(np.flatnonzero(~np.isin(mesh.ecell, mesh.becell).any(axis=1))[0])
    mesh.edge[k] = mesh.edge[k, ::-1].copy()
    c = M.flow_constants().as_array()
    st = _free_stream_residual(mesh, c)
    interior = np.ones(mesh.ncell, bool)
    interior[mesh.becell] = False
    scale = np.abs(c[6:]).max() * np.abs(mesh.x).max() * 1e-13
    assert np.abs(st.res[interior]).max() > 1e3 * scale


def test_res_calc_conserves():
    """Every edge adds +f to one cell and -f to the other: the residual sums
    to zero over the mesh up to rounding, for any flow state."""
    mesh = M.o_mesh(40, 20, "random", seed=5)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(11))
    st = O.AirfoilState(mesh, c, "f64")
    st.q *= 1.0 + 0.05 * rng.random(st.q.shape)
    O.adt_calc("f64", mesh, st.x, st.q, st.adt, c)
    O.res_calc("f64", mesh, st.x, st.q, st.adt, st.res, c)
    total = st.res.sum(axis=0)
    assert np.abs(total).max() < 1e-12 * np.abs(st.res).sum()
    assert np.abs(st.res).max() > 1e-3


def test_adt_calc_rectangle():
    """adt of an axis-aligned a x b rectangle with velocity (u, v) and sound
    speed c is (2|u| b + 2|v| a + 2 c (a + b)) / cfl."""
    a, b = 0.3, 0.7
    x = np.array([[0, 0], [a, 0], [a, b], [0, b]], np.float64)
    cell = np.array([[0, 1, 2, 3]], np.int32)
    mesh = M.AirfoilMesh(x=x, cell=cell, edge=np.zeros((0, 2), np.int32),
                         ecell=np.zeros((0, 2), np.int32), bedge=np.zeros((0, 2), np.int32),
                         becell=np.zeros(0, np.int32), bound=np.zeros(0, np.int32))
    k = M.flow_constants()
    c = k.as_array()
    q = np.array([[1.2, 0.3, -0.2, 2.9]])
    adt = np.zeros(1)
    O.adt_calc("f64", mesh, x, q, adt, c)
    u, v = q[0, 1] / q[0, 0], q[0, 2] / q[0, 0]
    snd = math.sqrt(k.gam * k.gm1 * (q[0, 3] / q[0, 0] - 0.5 * (u * u + v * v)))
    expect = (2 * abs(u) * b + 2 * abs(v) * a + 2 * snd * (a + b)) / k.cfl
    assert abs(adt[0] - expect) < 1e-14 * expect


def test_update_telescopes():
    """update moves q to qold - res/adt and zeroes res; rms is the sum of the
    squared moves (checked through q, not by repeating the formula)."""
    rng = np.random.Generator(np.random.PCG64(2))
    n = 1000
    qold = 1 + rng.random((n, 4))
    res = rng.standard_normal((n, 4))
    res0 = res.copy()
    adt = 0.5 + rng.random(n)
    q = np.zeros_like(qold)
    rms = O.update("f64", qold, q, res, adt)
    assert not res.any()
    assert np.allclose((qold - q) * adt[:, None], res0, rtol=1e-12, atol=1e-14)
    assert abs(rms - ((qold - q) ** 2).sum()) < 1e-12 * rms


def _perturbed_state(mesh, c, seed):
    """float32 x, q (free stream +-2%), adt > 0 and res of one mesh."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.ascontiguousarray(mesh.x.astype(np.float32))
    q = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    adt = (0.5 + rng.random(mesh.ncell)) * 10.0
    res = 1e-3 * rng.standard_normal((mesh.ncell, 4))
    f32 = lambda a: np.ascontiguousarray(a.astype(np.float32))
    return x, f32(q), f32(adt), f32(res)


def test_mixed_form_flux_loops_are_the_f32_loops():
    """adt_calc, res_calc and bres_calc of the precision form are the
    reference's own float loops (airfoil/{adt_calc,res_calc,bres_calc}.h in
    float): bit for bit the f32 oracle, which the reference build pins
    above.  A flux term evaluated in double would change the bits."""
    mesh = M.o_mesh(48, 24, "random", seed=4)
    c = M.flow_constants().as_array()
    x, q, adt0, res0 = _perturbed_state(mesh, c, 11)
    out = {}
    for form in ("f32", "mix"):
        adt = np.zeros(mesh.ncell, np.float32)
        O.adt_calc(form, mesh, x, q, adt, c)
        res = res0.copy()
        O.res_calc(form, mesh, x, q, adt0, res, c)
        bres = res.copy()
        O.bres_calc(form, mesh, x, q, adt0, bres, c)
        out[form] = (adt, res, bres)
    for a, b in zip(out["f32"], out["mix"]):
        assert a.dtype == np.float32 and np.array_equal(a, b)
    # the float loops differ from the double ones (so the pin is not vacuous)
    adt64 = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, x.astype(np.float64), q.astype(np.float64), adt64, c)
    assert not np.array_equal(adt64.astype(np.float32), out["mix"][0])


def test_mixed_form_update_is_the_rounded_f64_update():
    """update of the precision form: adti = 1/adt, del = adti res and
    qold - del formed in double, q rounded once to float, rms summed in
    double (airfoil/update.h:1-12).  Its q is float32 of the f64 update of
    the same (exactly upcast) inputs and its rms is that update's rms, bit
    for bit; an update carried out in float gives other bits."""
    mesh = M.o_mesh(48, 24, "random", seed=4)
    c = M.flow_constants().as_array()
    _, qold, adt, res = _perturbed_state(mesh, c, 12)
    res = np.ascontiguousarray(res * 50)  # moves of a few percent of q
    q_mix = np.zeros_like(qold)
    r_mix = res.copy()
    rms_mix = O.update("mix", qold, q_mix, r_mix, adt)
    q64 = np.zeros(qold.shape)
    r64 = res.astype(np.float64)
    rms64 = O.update("f64", qold.astype(np.float64), q64, r64, adt.astype(np.float64))
    assert np.array_equal(q_mix, q64.astype(np.float32))
    assert rms_mix == rms64
    assert not r_mix.any()
    q32 = np.zeros_like(qold)
    rms32 = O.update("f32", qold, q32, res.copy(), adt)
    assert np.count_nonzero(q32 != q_mix) > mesh.ncell // 10
    assert rms32 != rms_mix


def test_mixed_form_jacobi_update():
    """jac1 update of the precision form: u += du + alpha r in float (the
    f32 loop's bits) and u_sum, u_max accumulated in double: with du = r = 0
    (u unchanged, exactly) the sums equal the f64 update's of the upcast u,
    bit for bit, and differ from the float sums."""
    rng = np.random.Generator(np.random.PCG64(13))
    n = 4096
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    r, du, u = f32(rng.random(n)), f32(0.1 * rng.standard_normal(n)), f32(rng.random(n))
    u_mix, u_32 = u.copy(), u.copy()
    O.jac_update("mix", r, du.copy(), u_mix, 0.25)
    O.jac_update("f32", r, du.copy(), u_32, 0.25)
    assert np.array_equal(u_mix, u_32)
    zero = np.zeros(n, np.float32)
    s_mix = O.jac_update("mix", zero, zero.copy(), u_mix.copy(), 0.25)
    s_64 = O.jac_update("f64", zero.astype(np.float64), np.zeros(n), u_mix.astype(np.float64),
                        0.25)
    s_32 = O.jac_update("f32", zero, zero.copy(), u_mix.copy(), 0.25)
    assert s_mix == s_64
    assert s_mix[0] != s_32[0]


def test_mixed_form_close_to_f64():
    """The precision form (float state, double update) stays within float
    rounding of the double oracle over a short run."""
    mesh = M.o_mesh(48, 24, "random", seed=3)
    c = M.flow_constants().as_array()
    _, h64 = O.airfoil_seq(mesh, c, 50, "f64")
    _, hmx = O.airfoil_seq(mesh, c, 50, "mix")
    assert np.abs(hmx - h64).max() / np.abs(h64).max() < 1e-4


def test_mesh_batch_union_is_disjoint_meshes():
    """The union mesh of a batch marches each of its meshes exactly as alone
    (same free stream for all, so the plain airfoil loops apply): a wrong
    node, cell or edge offset in concat_meshes couples the meshes."""
    meshes = [M.o_mesh(16, 8, "random", seed=s) for s in (1, 2, 3)]
    c = M.flow_constants().as_array()
    batch = M.concat_meshes(meshes, [c] * 3)
    assert np.all(np.diff(batch.cmesh) >= 0) and np.all(np.diff(batch.bmesh) >= 0)
    assert batch.cell_off[-1] == batch.union.ncell
    st, h = O.airfoil_seq(batch.union, c, 5, "f64")
    sums = np.zeros(5)
    for i, m in enumerate(meshes):
        sti, hi = O.airfoil_seq(m, c, 5, "f64")
        q = st.q.reshape(-1, 4)[batch.cell_off[i]:batch.cell_off[i + 1]]
        assert np.array_equal(q, sti.q.reshape(-1, 4))
        sums += hi ** 2 * m.ncell
    np.testing.assert_allclose(h, np.sqrt(sums / batch.union.ncell), rtol=1e-13)
