"""Surrogate-screened evolution on the HIP path: the surrogate's forward and Adam kernels against the
oracle (oracle/surrogate.py, float64), child fingerprints against the oracle's probe play, and the
screened generation step's invariants, determinism and exact agreement with the plain step when
screening never engages."""
import numpy as np
import pytest

import surrogate as sgo

pytestmark = pytest.mark.gpu


def _runtime(pkg, wl, brood=4, seed=7):
    rtm = pkg.runtime
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=(brood + 2) * n * stride, generation_bytes=brood * n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=seed, variance_scale=1.0, tail_std=0.1)
    return rt


@pytest.fixture(scope="module")
def small(pkg):
    """4x4 grid on the 512-action catalog (fingerprints are compared with the oracle genome by genome)."""
    wl = pkg.synth.workload("grid16_a512")
    side = 4
    sub = dict(wl)
    sub["side"] = side
    sub["offsets"] = wl["offsets"][: side * side + 1]
    sub["ids"] = wl["ids"][: int(sub["offsets"][-1])]
    rt = _runtime(pkg, sub)
    rt.enable_screening(pkg.runtime.default_screen_config())
    yield sub, rt
    rt.close()


def test_surrogate_forward_matches_oracle(pkg, small):
    _, rt = small
    rng = np.random.default_rng(11)
    params = (rng.standard_normal(sgo.N_PARAMS) * 0.1).astype(np.float32)
    rt.set_surrogate_params(params)
    assert np.array_equal(rt.surrogate_params(), params)
    x = rng.standard_normal((37, 192)).astype(np.float32)  # not a multiple of the 8-row workgroup
    y = rt.surrogate_predict(x)
    ref, _, _ = sgo.forward(params, x)
    assert np.max(np.abs(y - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


def test_surrogate_adam_steps_match_oracle(pkg, small):
    _, rt = small
    rng = np.random.default_rng(12)
    params = (rng.standard_normal(sgo.N_PARAMS) * 0.1).astype(np.float32)
    rt.set_surrogate_params(params)  # also zeroes the Adam state and its step count
    ref = params.astype(np.float64)
    m = np.zeros(sgo.N_PARAMS)
    v = np.zeros(sgo.N_PARAMS)
    lr = pkg.runtime.default_screen_config().learning_rate
    # Adam's step g / (|g| + 1e-8) is ill-conditioned where a batch gradient cancels to ~1e-8: there the
    # fp32 sums' rounding (~1e-9) moves the step by a visible fraction of lr.  Such parameters (oracle
    # |g| < 1e-6 at some step; a handful of 33,025) are held to |error| <= lr, all others to 2e-6.
    loose = np.zeros(sgo.N_PARAMS, bool)
    for step in (1, 2, 3):
        x = rng.standard_normal((100, 192)).astype(np.float32)
        t = rng.standard_normal(100).astype(np.float32)
        rt.surrogate_train_step(x, t)
        g = sgo.gradient(ref, x, t)
        loose |= (np.abs(g) < 1e-6) & (g != 0)
        ref, m, v = sgo.adam_step(ref, m, v, step, x, t, lr)
        got = rt.surrogate_params()
        err = np.abs(got - ref)
        # an Adam update moves a parameter by about lr; a wrong sign, index or bias correction moves it by
        # that much, fp32 rounding by far less
        assert np.max(err[~loose]) <= 2e-6, step
        assert np.max(err[loose], initial=0.0) <= lr, step
        assert loose.sum() <= sgo.N_PARAMS // 1000
        assert np.max(np.abs(got - params)) > 0.5 * lr


@pytest.mark.parametrize("precision", ["fp32", "bf16_trunk"])
def test_fingerprints_match_oracle(pkg, orc, small, precision):
    """Probe play of three children against the oracle's fp64 policy on every probe whose top-2 margin is
    certain; "bf16_trunk": the fingerprint kernel of the BF16-trunk form (split-BF16 trunk layers 1 and 2,
    ~1e-5 relative from the FP32 trunk, far inside the margin)."""
    wl, rt = small
    rtm = pkg.runtime
    rt.set_eval_precision(rtm.EVAL_BF16_TRUNK if precision == "bf16_trunk" else rtm.EVAL_FP32)
    slots, _, _, _ = rt.generation(0)
    probes = rt.probe_states()
    assert probes.shape == (16, 4)
    catalog = wl["catalog"]
    A = wl["actions"]
    for k in range(16):  # the three guesses of a board are distinct words, none the secret
        words = {bytes(catalog[i]) for i in probes[k]}
        assert len(words) == 4 and probes[k].max() < A
    pf = np.array([[0.03 * (i + 1), 0.5 / (i + 2)] for i in range(len(slots))], np.float32)
    fp = rt.fingerprint(slots, pf)
    assert fp.shape == (len(slots), 192)
    checked = 0
    for i in (0, 5, 15):
        genome = rt.download_slot(slots[i])
        ref, certain = sgo.probe_features(genome, catalog, A, probes, pf[i])
        for p in range(64):
            d, k = divmod(p, 16)
            g = fp[i, 3 * p:3 * p + 3]
            # invariants whatever the guess
            assert g[0] in np.arange(11) / 10.0 or abs(g[0] * 10 - round(g[0] * 10)) < 1e-6
            if d == 0:
                assert g[1] == np.float32(10.0 * pf[i][k % 2])
            else:
                assert g[1] in (0.0, 1.0)
            assert g[2] in (0.0, 1.0) and (g[2] == 0.0 or g[0] == 1.0)
            if certain[p]:
                np.testing.assert_allclose(g, ref[3 * p:3 * p + 3], rtol=0, atol=1e-6, err_msg=f"genome {i} probe {p}")
                checked += 1
    assert checked >= 150
    rt.set_eval_precision(rtm.EVAL_FP32)


def _screened_run(pkg, wl, gens, screen_kw, seed=7):
    rtm = pkg.runtime
    rt = _runtime(pkg, wl, seed=seed)
    rt.enable_screening(rtm.default_screen_config(**screen_kw))
    counts = rtm.word_counts(wl["actions"], wl["actions"])
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    reports, fits = [], []
    for g in range(gens):
        rep = rt.advance_screened(9 + g, counts, cfg)
        info = rt.info()
        slots, fit, _, has = rt.generation(0)
        n = wl["side"] ** 2
        assert info.current_generation_size == n and has.all()
        assert info.free_slot_count == info.slot_count - n, "every unselected brood member is freed"
        assert len(set(slots.tolist())) == n
        reports.append((rep.screened, rep.brood, rep.evaluated, rep.audited, rep.best_fitness, rep.average_fitness))
        fits.append(fit.copy())
    return rt, reports, fits


def test_screened_steps_invariants_and_determinism(pkg):
    wl = pkg.synth.workload("grid16_a512")
    kw = dict(min_replay=256, spearman_threshold=-1.0)  # screening from the second step on
    rt, reps, fits = _screened_run(pkg, wl, 5, kw)
    n = 256
    assert reps[0][0] == 0 and reps[0][1] == n and reps[0][2] == n  # warm-up: plain
    for r in reps[1:]:
        assert r[0] == 1 and r[1] == 4 * n and n <= r[2] <= n + r[3]
    # the fitness a screened step reports is that of a full evaluation of the selected children
    counts = pkg.runtime.word_counts(512, 512)
    rt.evaluate(counts)
    _, fit2, _, _ = rt.generation(0)
    assert np.array_equal(fit2.view(np.uint32), fits[-1].view(np.uint32))
    rt.close()
    rt2, reps2, fits2 = _screened_run(pkg, wl, 5, kw)
    rt2.close()
    assert reps2 == reps
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(fits, fits2))


def test_unscreened_steps_equal_the_plain_step(pkg):
    """With screening never engaged (one child per cell) the screened step is the plain generation step,
    plus the evaluation of the children: same genomes, same slots, same fitness."""
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    gens = 3
    rt_s, reps, fits = _screened_run(pkg, wl, gens, dict(min_replay=1 << 30))
    assert all(r[0] == 0 for r in reps)
    rt_p = _runtime(pkg, wl)
    counts = rtm.word_counts(512, 512)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    for g in range(gens):
        rt_p.advance(9 + g, counts, cfg)
    rt_p.evaluate(counts)
    sp, fp_, _, _ = rt_p.generation(0)
    ss, fs, _, _ = rt_s.generation(0)
    assert np.array_equal(sp, ss)
    assert np.array_equal(fp_.view(np.uint32), fs.view(np.uint32))
    for i in (0, 100, 255):
        assert np.array_equal(rt_p.download_slot(sp[i]).view(np.uint32), rt_s.download_slot(ss[i]).view(np.uint32))
    rt_s.close()
    rt_p.close()


def test_two_phase_brood_children_equal_one_pass_breeding(pkg, orc):
    """A screened step breeds every brood member's policy and probe-vocabulary rows first and completes only
    the evaluated ones; each cell's selected child must equal the oracle's one-pass breeding of one of the
    cell's brood members (700 actions: 188 rows past the 512-word probe vocabulary are completed later)."""
    rtm = pkg.runtime
    A, side = 700, 8
    cat = pkg.synth.make_catalog(A)
    sec = pkg.synth.sample_secrets(A, 32)
    off, ids = pkg.synth.block_csr(side, side, 4, sec)
    wl = {"actions": A, "side": side, "catalog": cat, "offsets": off, "ids": ids}
    rt = _runtime(pkg, wl, seed=5)
    # a large audit slice so the 64-cell generation yields a finite rank correlation and screening starts
    rt.enable_screening(rtm.default_screen_config(min_replay=1, spearman_threshold=-1.0, audit_fraction=0.5))
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    assert rt.advance_screened(101, counts, cfg).screened == 0
    slots, fit, _, has = rt.generation(0)
    n = side * side
    n_vals = orc.policy_param_count() + 38 * A
    parents = [rt.download_slot(s)[:n_vals] for s in slots]
    seed = 202
    rep = rt.advance_screened(seed, counts, cfg)
    assert rep.screened == 1 and rep.brood == 4 * n
    children, _, _, _ = rt.generation(0)
    st, pairs = orc.plan_brood(fit, has, n, 4, side, 1, 2, orc.rank_weights(0.5, max(n, 25)), seed)
    assert st == 0
    st, plain = orc.plan(fit, has, n, n, side, 1, 2, orc.rank_weights(0.5, max(n, 25)), seed)
    assert st == 0 and np.array_equal(pairs[:n], plain), "brood member 0 is the plain plan's child"
    temps = (cfg.crossover_temperature_level1, cfg.crossover_temperature_level2, cfg.crossover_temperature_level3)
    picked = []
    for c in range(n):
        got = rt.download_slot(children[c])
        assert not got[n_vals:].any()
        match = [b for b in range(4)
                 if np.array_equal(got[:n_vals].view(np.uint32), orc.breed_child(
                     parents[pairs[b * n + c][0]], parents[pairs[b * n + c][1]], A, A, seed, 1, b * n + c, *temps,
                     cfg.mutation_probability, cfg.mutation_sigma, 0.0)[0].view(np.uint32))]
        assert match, c
        picked.append(match[0])
    assert len(set(picked)) > 1, "the surrogate picks different brood members in different cells"
    rt.close()


def test_graph_replayed_training_equals_eager_launches(pkg):
    """A generation's training steps replay as one captured hipGraph; over several generations (eager work
    and synchronisations between the replays) every replay must leave what eager launches leave."""
    wl = pkg.synth.workload("grid16_a512")
    kw = dict(min_replay=256, spearman_threshold=-1.0)
    rt_g, reps_g, fits_g = _screened_run(pkg, wl, 5, kw)
    rt_e, reps_e, fits_e = _screened_run(pkg, wl, 5, dict(kw, flags=pkg.runtime.SCREEN_EAGER_TRAINING))
    assert reps_g == reps_e
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(fits_g, fits_e))
    assert np.array_equal(rt_g.surrogate_params().view(np.uint32), rt_e.surrogate_params().view(np.uint32))
    rt_g.close()
    rt_e.close()


def test_screened_step_reevaluates_after_new_secrets(pkg):
    """A screened step plans from the current generation's fitness.  When the secrets change between steps
    (here: a different per-cell CSR), that fitness must be measured again on the new secrets, as the
    reference evaluates at every generation boundary (slab_runtime.cu:1374-1424): a runtime that re-uploads
    and steps must end exactly like one that also evaluates explicitly before stepping."""
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    kw = dict(min_replay=256, spearman_threshold=-1.0)
    counts = rtm.word_counts(512, 512)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    # the same block shards with the secret lists rotated within each cell: another CSR of equal shape
    offs = wl["offsets"]
    ids2 = np.concatenate([np.roll(wl["ids"][offs[c]:offs[c + 1]], 3) for c in range(len(offs) - 1)])
    ids2 = (ids2 + 1) % 512
    runs = []
    for explicit in (False, True):
        rt = _runtime(pkg, wl)
        rt.enable_screening(rtm.default_screen_config(**kw))
        for g in range(2):
            rt.advance_screened(9 + g, counts, cfg)
        rt.upload_secrets_csr(offs, ids2.astype(np.uint32))
        if explicit:
            rt.evaluate(counts)
        stale = rt.generation(0)[1].copy()
        rt.advance_screened(11, counts, cfg)
        pairs = rt.plan().copy()
        slots, fit, _, _ = rt.generation(0)
        runs.append((stale, pairs, slots.copy(), fit.copy()))
        rt.close()
    (s0, p0, sl0, f0), (s1, p1, sl1, f1) = runs
    assert not np.array_equal(s0, s1), "the new secrets change the current generation's fitness"
    assert np.array_equal(p0, p1)
    assert np.array_equal(sl0, sl1)
    assert np.array_equal(f0.view(np.uint32), f1.view(np.uint32))
