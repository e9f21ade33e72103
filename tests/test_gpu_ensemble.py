"""GPU parity of the batched ensemble (persistent work-queue force kernel,
size-bucket tuning, shared-dt Hermite steps, hipGraph replay) against the
oracle run system by system, through the C ABI."""
import numpy as np
import pytest

from conftest import body_relerr

pytestmark = pytest.mark.gpu

EDGE_SIZES = [1, 2, 3, 63, 64, 65, 128, 255, 256, 257, 700, 1025]


def _banner(msg):
    print(f"[gpu-ensemble] {msg}", flush=True)


@pytest.fixture(scope="module")
def small(nb):
    counts = np.concatenate([EDGE_SIZES, nb.ensemble_sizes(28, 4, 400, 7)]).astype(np.int32)
    eps2 = (1.0 / (4.0 * counts.astype(float) ** (1 / 3))) ** 2
    state = nb.ensemble_state(counts, 11)
    return counts, eps2, state


def check_systems(ens, f, state, oracle, eps2, tol_max=1e-12, mixed=False):
    worst = 0.0
    meds = []
    for k in range(len(ens.counts)):
        s = ens.system(state, k)
        ref = oracle.force(s, eps2[k])
        fk = ens.system(f, k)
        if s.shape[1] == 1:
            assert np.all(fk == 0.0)
            continue
        ea = body_relerr(fk[:3], ref[:3])
        if mixed:
            meds.append(np.median(ea))
        else:
            ej = body_relerr(fk[3:], ref[3:])
            worst = max(worst, ea.max(), ej.max())
    if mixed:
        assert np.median(meds) <= 1e-6 and max(meds) <= 1e-5, (np.median(meds), max(meds))
    else:
        assert worst <= tol_max, worst
    return worst


def test_initial_forces_fp64(nb, oracle, small):
    counts, eps2, state = small
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        _, f = e.download()
        w = check_systems(e, f, state, oracle, eps2)
        _banner(f"fp64 ensemble of {len(counts)} systems: worst per-body rel err {w:.2e}")


def test_initial_forces_mixed(nb, oracle, small):
    counts, eps2, state = small
    with nb.Backend("mixed") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        _, f = e.download()
        check_systems(e, f, state, oracle, eps2, mixed=True)


@pytest.mark.parametrize("ipl", [1, 2])
@pytest.mark.parametrize("tile", [64, 128, 256])
@pytest.mark.parametrize("split", [1, 2, 4, 8])
def test_tunings_agree(nb, oracle, small, tile, split, ipl):
    counts, eps2, state = small
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.set_tuning([tile] * 4, [split] * 4)
        e.set_targets_per_lane(ipl)
        assert e.tuning() == {"tile": [tile] * 4, "jsplit": [split] * 4, "ipl": ipl}
        e.upload(state)
        _, f = e.download()
        check_systems(e, f, state, oracle, eps2)


@pytest.mark.parametrize("split", [1, 4])
def test_targets_per_lane_do_not_change_sums(nb, small, split):
    """Two targets per lane share each source record but keep each target's
    summation order, so FP64 and mixed results are bit-identical to one."""
    counts, eps2, state = small
    for prec in ("fp64", "mixed"):
        outs = []
        with nb.Backend(prec) as b, nb.Ensemble(b, counts, eps2) as e:
            e.set_tuning([128] * 4, [split] * 4)
            for ipl in (1, 2, 1):
                e.set_targets_per_lane(ipl)
                e.upload(state)
                outs.append(e.download()[1])
        assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_tile_length_does_not_change_sums(nb, small):
    """With no j-split every target sums its sources in index order, so the
    LDS tile length cannot change the FP64 or mixed result."""
    counts, eps2, state = small
    for prec in ("fp64", "mixed"):
        outs = []
        with nb.Backend(prec) as b, nb.Ensemble(b, counts, eps2) as e:
            for tile in (64, 128, 256):
                e.set_tuning([tile] * 4, [1] * 4)
                e.upload(state)
                outs.append(e.download()[1])
        assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_steps_vs_oracle(nb, oracle, small):
    counts, eps2, state = small
    dt, nsteps = 1 / 256, 4
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.set_tuning([64, 128, 128, 256], [1, 2, 2, 4])
        e.set_targets_per_lane(2)
        e.upload(state)
        e.set_dt(dt)
        e.step(nsteps)
        st, f = e.download()
    worst = 0.0
    for k in range(len(counts)):
        o = oracle.run(e.system(state, k), eps2[k], dt, nsteps)
        worst = max(worst, np.abs(e.system(st, k) - o["state"]).max())
    _banner(f"4 ensemble steps: worst state difference {worst:.2e}")
    assert worst <= 1e-12


def test_graph_matches_eager(nb, small):
    """Every replay of the captured 16-step ensemble graph is checked against
    an eager twin, with eager energy evaluations between replays."""
    counts, eps2, state = small
    with nb.Backend("mixed") as bg, nb.Backend("mixed") as be, \
            nb.Ensemble(bg, counts, eps2) as g, nb.Ensemble(be, counts, eps2) as e:
        for x in (g, e):
            x.set_tuning([64, 128, 128, 256], [1, 2, 4, 8])
            x.upload(state)
            x.set_dt(1 / 256)
        for r in range(4):
            g.step(16, graph=True)
            e.step(16)
            kg, ug = g.energy()
            ke, ue = e.energy()
            assert np.array_equal(kg, ke) and np.array_equal(ug, ue)
            sg, fg = g.download()
            se, fe = e.download()
            assert np.array_equal(sg, se) and np.array_equal(fg, fe), f"replay {r}"


def test_energy_per_system(nb, oracle, small):
    counts, eps2, state = small
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        kin, pot = e.energy()
    for k in range(len(counts)):
        s = e.system(state, k)
        assert kin[k] == pytest.approx(oracle.kinetic(s), rel=1e-13)
        if counts[k] > 1:
            assert pot[k] == pytest.approx(oracle.potential(s, eps2[k]), rel=1e-12)
        else:
            assert pot[k] == 0.0


def test_autotune_then_parity(nb, oracle, small):
    counts, eps2, state = small
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        t = e.autotune()
        _banner(f"autotune: {t}")
        assert all(x in (64, 128, 256) for x in t["tile"])
        assert all(x in (1, 2, 4, 8) for x in t["jsplit"])
        e.upload(state)
        _, f = e.download()
        check_systems(e, f, state, oracle, eps2)


def test_zero_softening_system(nb, oracle):
    counts = np.array([50, 130], dtype=np.int32)
    eps2 = np.array([0.0, 1e-3])
    state = nb.ensemble_state(counts, 5)
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        _, f = e.download()
        check_systems(e, f, state, oracle, eps2)


def test_bad_arguments(nb):
    with nb.Backend("fp64") as b:
        with pytest.raises(nb.NBodyError) as ex:
            nb.Ensemble(b, [3, 0, 4], 1e-3)
        assert ex.value.code == nb.ERR_ARG
        with nb.Ensemble(b, [3, 4], 1e-3) as e:
            with pytest.raises(nb.NBodyError) as ex:
                e.step(1)
            assert ex.value.code == nb.ERR_STATE
            with pytest.raises(nb.NBodyError) as ex:
                e.set_tuning([96, 64, 64, 64], [1, 1, 1, 1])
            assert ex.value.code == nb.ERR_ARG
            with pytest.raises(nb.NBodyError) as ex:
                e.set_targets_per_lane(3)
            assert ex.value.code == nb.ERR_ARG


# ---- symmetric pairing (context created with pairing="symmetric"): systems
# of >= 2 blocks (256 bodies in mixed precision, 512 in FP64) evaluate each
# unordered pair once; sizes with and without a partial last block, either
# side of both thresholds
SYM_SIZES = [1024, 1025, 1100, 1224, 1324, 1500, 2048, 3001, 600, 511, 512, 513, 767, 100, 5, 1]


@pytest.fixture(scope="module")
def symset(nb):
    counts = np.array(SYM_SIZES, dtype=np.int32)
    eps2 = (1.0 / (4.0 * counts.astype(float) ** (1 / 3))) ** 2
    state = nb.ensemble_state(counts, 23)
    return counts, eps2, state


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_ensemble_vs_oracle(nb, oracle, symset, precision):
    counts, eps2, state = symset
    with nb.Backend(precision, pairing="symmetric") as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        _, f = e.download()
        w = check_systems(e, f, state, oracle, eps2, mixed=(precision == "mixed"))
    _banner(f"symmetric {precision} ensemble {SYM_SIZES}: worst fp64 per-body rel err {w:.2e}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_ensemble_matches_directed_and_graph(nb, symset, precision):
    counts, eps2, state = symset
    out = {}
    for pairing, graph in (("directed", False), ("symmetric", False), ("symmetric", True)):
        with nb.Backend(precision, pairing=pairing) as b, nb.Ensemble(b, counts, eps2) as e:
            e.upload(state)
            e.autotune()
            e.set_dt(1 / 256)
            e.step(16, graph=graph)
            out[(pairing, graph)] = e.download()
    xd = out[("directed", False)][0]
    xs, xg = out[("symmetric", False)][0], out[("symmetric", True)][0]
    d = np.abs(xs[:3] - xd[:3]).max()
    _banner(f"{precision}: 16 steps, max |x_sym - x_dir| {d:.2e}, |graph - eager| {np.abs(xs - xg).max():.2e}")
    assert d <= (1e-12 if precision == "fp64" else 1e-6)
    assert np.abs(xs - xg).max() <= (1e-13 if precision == "fp64" else 1e-9)


@pytest.mark.parametrize("precision,pairing", [("fp64", "directed"), ("fp64", "symmetric"), ("mixed", "directed"),
                                              ("mixed", "symmetric")])
def test_full_cluster_set(nb, oracle, precision, pairing):
    """The 8192-cluster workload itself (n_k log-uniform in [8, 4096], seed 4,
    eps_k = 1/(4 n_k^(1/3))): the momentum sum of every system vanishes to
    rounding, and a sample of systems (the six largest, which the symmetric
    kernel covers, the smallest, random ones) matches the oracle to the
    precision's bar."""
    counts = nb.ensemble_sizes(8192, 8, 4096, 4)
    eps2 = (1.0 / (4.0 * counts.astype(float) ** (1.0 / 3.0))) ** 2
    state = nb.ensemble_state(counts, 4)
    _banner(f"full cluster set {precision} {pairing}: {len(counts)} systems, {counts.sum()} bodies")
    with nb.Backend(precision, pairing=pairing) as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(state)
        _, f = e.download()
        rel = []
        for k in range(len(counts)):
            s, fk = e.system(state, k), e.system(f, k)
            m = s[3]
            for c in range(3):
                rel.append(abs(np.sum(m * fk[c])) / max(np.sum(np.abs(m * fk[c])), 1e-300))
        pick = np.unique(np.concatenate([np.argsort(counts)[-6:], [np.argmin(counts)],
                                         np.random.default_rng(8).choice(len(counts), 9, replace=False)]))
        worst, meds = 0.0, []
        for k in pick:
            s, fk = e.system(state, k), e.system(f, k)
            ref = oracle.force(s, eps2[k], nthreads=8)
            ea = body_relerr(fk[:3], ref[:3])
            ej = body_relerr(fk[3:], ref[3:])
            worst = max(worst, ea.max(), ej.max())
            meds.append(np.median(ea))
    _banner(f"  momentum sum worst {max(rel):.1e}; sample of {len(pick)} systems (largest {counts.max()}): "
            f"worst {worst:.2e}, median of medians {np.median(meds):.2e}")
    assert max(rel) <= (1e-12 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        assert worst <= 1e-12
    else:
        assert np.median(meds) <= 1e-6 and max(meds) <= 1e-5


@pytest.mark.parametrize("pairing", ["directed", "symmetric"])
def test_mixed_systems_moving_apart(nb, oracle, symset, pairing):
    """Every system of the symmetric set moving at its own bulk velocity of
    ~1e3 (each system's velocities are taken relative to its own
    centre-of-mass velocity before they are rounded to FP32): per-system
    median acceleration and jerk errors against the oracle meet the mixed
    bars as at rest (median over systems <= 1e-6, worst system <= 1e-5; FP32
    rounded velocities would put the jerk near 2e-5 everywhere)."""
    counts, eps2, state = symset
    st = state.copy()
    boost = np.random.default_rng(41).uniform(-1e3, 1e3, (len(counts), 3))
    for k in range(len(counts)):
        o = int(np.sum(counts[:k]))
        st[4:7, o:o + counts[k]] += boost[k][:, None]
    with nb.Backend("mixed", pairing=pairing) as b, nb.Ensemble(b, counts, eps2) as e:
        e.upload(st)
        _, f = e.download()
        meds = []
        for k in range(len(counts)):
            if counts[k] < 2:
                continue
            s, fk = e.system(st, k), e.system(f, k)
            ref = oracle.force(s, eps2[k], nthreads=8)
            meds.append((np.median(body_relerr(fk[:3], ref[:3])), np.median(body_relerr(fk[3:], ref[3:]))))
    meds = np.array(meds)
    _banner(f"moving systems, mixed {pairing}: worst per-system median acc {meds[:, 0].max():.2e}, "
            f"jerk {meds[:, 1].max():.2e}")
    assert np.median(meds[:, 0]) <= 1e-6 and np.median(meds[:, 1]) <= 1e-6
    assert meds[:, 0].max() <= 1e-5 and meds[:, 1].max() <= 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scalar_load_kernel_bitwise_equals_lds_tiles(nb, oracle, precision):
    """Tile 0 puts a bucket's systems on the scalar-load kernel (256 targets
    of one system per four-wave workgroup, sources read with scalar loads):
    the same sums, bit for bit, as the persistent LDS-tile kernel with one
    split, for the large bucket alone and for every bucket; and against the
    oracle.  A system's last workgroup then holds 1..256 targets, its waves
    past the last one leaving at once: 128/129, 192/193, 255 and 257 put that
    edge on every wave boundary."""
    counts = np.array([1, 2, 63, 64, 65, 128, 129, 192, 193, 255, 257, 300, 1025, 1100, 2048, 2049, 3001, 4096],
                      dtype=np.int32)
    eps2 = (1.0 / (4.0 * counts.astype(float) ** (1 / 3))) ** 2
    state = nb.ensemble_state(counts, 29)
    out = {}
    for name, tiles in (("lds", [256] * 4), ("large", [256, 256, 256, 0]), ("all", [0] * 4)):
        with nb.Backend(precision) as b, nb.Ensemble(b, counts, eps2) as e:
            e.set_tuning(tiles, [1] * 4)
            e.upload(state)
            out[name] = e.download()[1]
            if name == "large":
                check_systems(e, out[name], state, oracle, eps2, mixed=(precision == "mixed"))
    _banner(f"{precision} scalar-load ensemble kernel vs LDS tiles")
    assert np.array_equal(out["large"], out["lds"])
    assert np.array_equal(out["all"], out["lds"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_context_with_scalar_load_buckets(nb, oracle, symset, precision):
    """Regression: with symmetric pairing, a bucket set to the scalar-load
    kernel (tile 0) still sends its symmetric systems' diagonal blocks to the
    persistent LDS kernel, which needs a tile length (it takes 256); the
    forces match the oracle."""
    counts, eps2, state = symset
    with nb.Backend(precision, pairing="symmetric") as b, nb.Ensemble(b, counts, eps2) as e:
        e.set_tuning([0, 0, 0, 0], [1, 1, 1, 1])
        e.upload(state)
        _, f = e.download()
        check_systems(e, f, state, oracle, eps2, mixed=(precision == "mixed"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_stream_step_graph_matches_eager(nb, small, precision):
    """The largest bucket on the scalar-load kernel and the others on the
    persistent kernel: the two launches run on two streams, forked and joined
    with events, and a captured step graph holds both branches.  Sixteen
    steps replayed from the graph equal sixteen eager steps bit for bit."""
    counts, eps2, state = small
    out = {}
    for graph in (False, True):
        with nb.Backend(precision) as b, nb.Ensemble(b, counts, eps2) as e:
            e.set_tuning([64, 128, 256, 0], [1, 2, 1, 1])
            e.upload(state)
            e.set_dt(1.0 / 256)
            e.step(16, graph=graph)
            out[graph] = e.download()
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_step_paths_bitwise(nb, small, precision):
    """A step corrects and predicts the scalar-load kernel's systems inside
    that kernel (double-buffered predictions and records) and every other
    system in a separate one-thread-per-body pass; the arithmetic is the
    same, so five steps (odd: the run ends on the other prediction buffer)
    give the same states and forces bit for bit with every bucket on the LDS
    kernel, every bucket on the scalar-load kernel, or a mixture.  Steps
    split across calls, eager and graph-replayed (a captured sequence
    started from either buffer), equal the same number of eager steps."""
    counts, eps2, state = small
    dt = 1.0 / 256

    def run(tiles, plan):
        with nb.Backend(precision) as b, nb.Ensemble(b, counts, eps2) as e:
            e.set_tuning(tiles, [1, 1, 1, 1])
            e.upload(state)
            e.set_dt(dt)
            for n, graph in plan:
                e.step(n, graph=graph)
            return e.download()

    five = [(5, False)]
    ref = run([256] * 4, five)
    for tiles in ([0] * 4, [64, 0, 256, 0]):
        got = run(tiles, five)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), tiles
    eager = run([64, 0, 256, 0], [(37, False)])
    mixed_plan = run([64, 0, 256, 0], [(3, False), (16, True), (1, False), (16, True), (1, True)])
    assert np.array_equal(mixed_plan[0], eager[0]) and np.array_equal(mixed_plan[1], eager[1])
    _banner(f"{precision} ensemble step paths and graph/eager splits bitwise equal")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zero_softening_steps(nb, oracle, precision):
    """Systems with eps2 = 0 take the kernels that drop the self pair; their
    steps advance the targets inside those kernels too.  Three steps with
    every bucket on the scalar-load kernel equal three steps on the LDS
    kernel bit for bit, and (FP64) the oracle's run of each system."""
    counts = np.array([50, 130, 300], dtype=np.int32)
    eps2 = np.array([0.0, 1e-3, 0.0])
    state = nb.ensemble_state(counts, 5)
    dt, nsteps = 1.0 / 256, 3
    out = {}
    for tiles in ([0] * 4, [64] * 4):
        with nb.Backend(precision) as b, nb.Ensemble(b, counts, eps2) as e:
            e.set_tuning(tiles, [1] * 4)
            e.upload(state)
            e.set_dt(dt)
            e.step(nsteps)
            out[tiles[0]] = e.download()
    assert np.array_equal(out[0][0], out[64][0]) and np.array_equal(out[0][1], out[64][1])
    if precision == "fp64":
        st = out[0][0]
        o = 0
        for k, n in enumerate(counts):
            ref = oracle.run(state[:, o:o + n], eps2[k], dt, nsteps)
            assert np.abs(st[:, o:o + n] - ref["state"]).max() <= 1e-12
            o += n
    _banner(f"{precision} zero-softening ensemble steps on both kernels")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_tilings_and_step_plans(nb, precision):
    """Seeded random ensembles (1..40 systems of 1..2000 bodies), random
    bucket tilings (LDS tiles 64/128/256 or the scalar-load kernel, one
    split) and random plans that split the steps between eager calls and
    graph replays of either length parity: every run equals the same number
    of eager steps on LDS tiles of 256, states and forces bit for bit."""
    rng = np.random.default_rng(1234 if precision == "fp64" else 4321)
    for case in range(6):
        nsys = int(rng.integers(1, 41))
        counts = rng.integers(1, 2001, nsys).astype(np.int32)
        eps2 = (1.0 / (4.0 * counts.astype(float) ** (1 / 3))) ** 2
        state = nb.ensemble_state(counts, 100 + case)
        tiles = [int(t) for t in rng.choice([0, 64, 128, 256], 4)]
        plan = [(int(rng.integers(1, 20)), bool(rng.integers(0, 2))) for _ in range(int(rng.integers(1, 5)))]
        total = sum(n for n, _ in plan)
        print(f"[gpu-ensemble] case {case}: {nsys} systems, tiles {tiles}, plan {plan}", flush=True)
        out = {}
        for name, tl, pl in (("ref", [256] * 4, [(total, False)]), ("run", tiles, plan)):
            with nb.Backend(precision) as b, nb.Ensemble(b, counts, eps2) as e:
                e.set_tuning(tl, [1] * 4)
                e.upload(state)
                e.set_dt(1.0 / 256)
                for n, graph in pl:
                    e.step(n, graph=graph)
                out[name] = e.download()
        assert np.array_equal(out["run"][0], out["ref"][0]), (case, tiles, plan)
        assert np.array_equal(out["run"][1], out["ref"][1]), (case, tiles, plan)


@pytest.mark.parametrize("tile", [64, 0])
def test_systems_bitwise_those_of_single_system_runs(nb, small, tile):
    """FP64, every bucket unsplit (LDS tiles of 64, or the scalar-load
    kernel): each target sums its sources in order from the system's first
    body, as the single-system LDS kernel does with one split, and the step
    uses the same predictor and corrector arithmetic.  So every system's
    initial forces, and its state and forces after 6 steps, equal those of
    the same system run alone, bit for bit."""
    counts, eps2, state = small
    dt, nsteps = 1.0 / 256, 6
    with nb.Backend("fp64") as b, nb.Ensemble(b, counts, eps2) as e:
        e.set_tuning([tile] * 4, [1] * 4)
        e.upload(state)
        _, f0 = e.download()
        e.set_dt(dt)
        e.step(nsteps)
        sT, fT = e.download()
    off = np.concatenate([[0], np.cumsum(counts)])
    for k in range(len(counts)):
        a, z = off[k], off[k + 1]
        with nb.Backend("fp64", eps2[k], block=64, jsplit=1, ipl=1, sources="lds") as b1:
            b1.upload(state[:, a:z])
            _, g0 = b1.download()
            b1.set_dt(dt)
            b1.step(nsteps)
            s1, g1 = b1.download()
        assert np.array_equal(f0[:, a:z], g0), f"system {k} (n={counts[k]}): initial forces"
        assert np.array_equal(sT[:, a:z], s1), f"system {k} (n={counts[k]}): state after {nsteps} steps"
        assert np.array_equal(fT[:, a:z], g1), f"system {k} (n={counts[k]}): forces after {nsteps} steps"
    _banner(f"tile {tile}: {len(counts)} systems bitwise equal to single-system runs over {nsteps} steps")
