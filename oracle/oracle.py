"""ctypes binding of the CPU oracle (liboracle.so) and of the reference-header probe (_ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, never by the product.  See ne_oracle.cpp for what each function restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libneref.so")

GAME_TRACE_DTYPE = np.dtype([("guesses", "<u4", 6), ("best", "<f4", 6), ("second", "<f4", 6),
                             ("decisions", "<u4"), ("turn_count", "<u4"), ("score", "<f4"), ("won", "<u4")])

P = C.c_void_p
U64, U32, I32, F32 = C.c_uint64, C.c_uint32, C.c_int, C.c_float

_SIGS = {
    "oro_philox": (None, [P, P, P]),
    "oro_u01": (F32, [U32]),
    "oro_normal_from_words": (F32, [U32, U32]),
    "oro_det_log": (C.c_double, [C.c_double]),
    "oro_det_sqrt": (C.c_double, [C.c_double]),
    "oro_policy_param_count": (U64, [I32]),
    "oro_feedback": (None, [P, P, P]),
    "oro_policy_forward": (I32, [P, I32, P, P, I32, P, P]),
    "oro_fixed_features": (None, [P, P]),
    "oro_policy_forward_t16": (I32, [P, I32, P, P, I32, P]),
    "oro_bf16_round": (None, [P, P, U64]),
    "oro_select_guess": (I32, [P, U64, P, P, P, I32, P, P, P, P, P, P]),
    "oro_select_guess_layout": (I32, [P, I32, U64, P, P, P, I32, P, P, P, P, P, P]),
    "oro_encode_turn": (None, [P, I32, P, P, P, P]),
    "oro_score_action": (F32, [P, P, P]),
    "oro_evaluate": (I32, [P, U64, I32, U64, U64, P, U64, P, P, I32, I32, P, P, P, P, P, P]),
    "oro_replay_game": (I32, [P, I32, U64, P, P, U64, U64, P, U32, P, P, P, P, P]),
    "oro_summary": (None, [P, U64, P, P, P]),
    "oro_normalize_fitness": (F32, [F32, U64]),
    "oro_rank_weights": (None, [F32, U64, P]),
    "oro_select_one": (I32, [P, P, P, U64, P, U64, F32, U64, P]),
    "oro_bernoulli": (I32, [F32, U32]),
    "oro_moore": (I32, [U64, U64, U64, U64, P]),
    "oro_chebyshev": (U64, [U64, U64, U64, U64]),
    "oro_plan": (I32, [P, P, U64, U64, U64, U64, U64, P, U32, P]),
    "oro_plan_brood": (I32, [P, P, U64, U64, U64, U64, U64, P, U32, P]),
    "oro_mt19937_64_nth": (U64, [U64, U64]),
    "oro_sample_plan": (I32, [U64, U64, U64, U64, U64, U64, P, P, P, P]),
    "oro_xover_decisions": (None, [I32, U64, U32, U32, U32, F32, F32, F32, P]),
    "oro_xover_record_words": (U64, []),
    "oro_assemble_slots": (I32, [U32, P, P, P, P, U32, P, U32, P, P, U32, P]),
    "oro_hint_group": (I32, [P, P, P, P]),
    "oro_repack_slots": (I32, [U32, P, U32, P, U32, P, P, P, P, P]),
    "oro_median_tail_norm": (F32, [P, U64]),
    "oro_inject": (I32, [P, P, I32, U64, P, U64, U64, I32]),
    "oro_breed_child": (None, [I32, P, P, P, U64, U64, U32, U32, U32, F32, F32, F32, F32, F32, F32, P]),
    "oro_bootstrap_genome": (None, [I32, P, U64, U32, U32, U32, P, F32]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle not built: run oracle/build_oracle.py")
        L = C.CDLL(ORACLE_LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref():
    """The reference-header probe, or None when it was not built (reference sources absent)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = C.CDLL(REF_LIB)
        R.ref_feedback.argtypes = [P, P, P]
        R.ref_feedback.restype = I32
        R.ref_hint_group.argtypes = [P, P, P, P]
        R.ref_hint_group.restype = I32
        R.ref_moore.argtypes = [U64, U64, U64, U64, P]
        R.ref_moore.restype = I32
        R.ref_chebyshev.argtypes = [U64, U64, U64, U64]
        R.ref_chebyshev.restype = U64
        R.ref_shard_csr.argtypes = [U64] * 10 + [P, U64, P, P, U64]
        R.ref_shard_csr.restype = I32
        R.ref_word_counts_image.argtypes = [P, P]
        R.ref_word_counts_image.restype = I32
        R.ref_assembly_image.argtypes = [U64, P, P]
        R.ref_assembly_image.restype = I32
        R.ref_assembly_defaults.argtypes = [P, P]
        R.ref_assembly_defaults.restype = None
        R.ref_grid_shape_image.argtypes = [U64, U64, P]
        R.ref_grid_shape_image.restype = I32
        R.ref_checkpoint_sizes.argtypes = [P]
        R.ref_checkpoint_sizes.restype = None
        _ref = R
    return _ref


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oro_philox(ptr(c), ptr(k), ptr(out))
    return out


def feedback(guess, sol):
    g = np.asarray(guess, np.uint8)
    s = np.asarray(sol, np.uint8)
    out = np.zeros(5, np.uint8)
    lib().oro_feedback(ptr(g), ptr(s), ptr(out))
    return out


def word(text):
    return np.frombuffer(text.encode(), np.uint8) - ord("A")


def policy_param_count(enc_hidden=0):
    return int(lib().oro_policy_param_count(enc_hidden))


def policy_forward(genome, turns, enc_hidden=0):
    """turns: list of (word letters[5], tiles[5]).  Returns (fp32 policy, fp64 policy)."""
    g = np.ascontiguousarray(genome, np.float32)
    tc = len(turns)
    letters = np.zeros((max(tc, 1), 5), np.uint8)
    tiles = np.zeros((max(tc, 1), 5), np.uint8)
    for i, (w, t) in enumerate(turns):
        letters[i] = w
        tiles[i] = t
    o32 = np.zeros(64, np.float32)
    o64 = np.zeros(64, np.float64)
    st = lib().oro_policy_forward(ptr(g), enc_hidden, ptr(letters), ptr(tiles), tc, ptr(o32), ptr(o64))
    if st != 0:
        raise RuntimeError(f"oro_policy_forward status {st}")
    return o32, o64


def policy_forward_t16(genome, turns, enc_hidden=0):
    """The BF16-trunk policy (evaluation mode 2) of a board of (word letters[5], tiles[5]) turns: fp32."""
    g = np.ascontiguousarray(genome, np.float32)
    tc = len(turns)
    letters = np.zeros((max(tc, 1), 5), np.uint8)
    tiles = np.zeros((max(tc, 1), 5), np.uint8)
    for i, (w, t) in enumerate(turns):
        letters[i] = w
        tiles[i] = t
    out = np.zeros(64, np.float32)
    st = lib().oro_policy_forward_t16(ptr(g), enc_hidden, ptr(letters), ptr(tiles), tc, ptr(out))
    if st != 0:
        raise RuntimeError(f"oro_policy_forward_t16 status {st}")
    return out


def bf16_round(x):
    """Round-to-nearest-even bfloat16 rounding of fp32 values, returned as fp32."""
    a = np.ascontiguousarray(x, np.float32).ravel()
    out = np.empty_like(a)
    lib().oro_bf16_round(ptr(a), ptr(out), a.size)
    return out.reshape(np.shape(x))


def select_guess(genome, A, catalog, turns, enc_hidden=0):
    """The next guess for a board of (word letters[5], tiles[5]) turns: dict with status, the fp32 choice
    (index, best, second) and the fp64 one.  genome: fp32 values in the layout of `enc_hidden`."""
    g = np.ascontiguousarray(genome, np.float32)
    cat = np.ascontiguousarray(catalog, np.uint8)
    tc = len(turns)
    letters = np.zeros((max(tc, 1), 5), np.uint8)
    tiles = np.zeros((max(tc, 1), 5), np.uint8)
    for i, (w, t) in enumerate(turns):
        letters[i] = w
        tiles[i] = t
    i32, i64 = C.c_uint64(), C.c_uint64()
    b32, s32 = C.c_float(), C.c_float()
    b64, s64 = C.c_double(), C.c_double()
    st = lib().oro_select_guess_layout(ptr(g), enc_hidden, A, ptr(cat), ptr(letters), ptr(tiles), tc, C.byref(i32),
                                       C.byref(b32), C.byref(s32), C.byref(i64), C.byref(b64), C.byref(s64))
    return dict(status=st, index=i32.value, best=np.float32(b32.value), second=np.float32(s32.value),
                index64=i64.value, best64=b64.value, second64=s64.value)


def evaluate(genomes, enc_hidden, A_genome, A_runtime, catalog, offsets, ids, mode=0, threads=1, trace=False):
    """genomes: list/array of fp32 genome vectors.  Returns dict with fitness, raw, local_w, decisions,
    traces (per game, ordered like the runtime's trace)."""
    N = len(genomes)
    keep = [np.ascontiguousarray(g, np.float32) for g in genomes]
    ptrs = (C.c_void_p * N)(*[g.ctypes.data for g in keep])
    cat = np.ascontiguousarray(catalog, np.uint8)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    ids = np.ascontiguousarray(ids, np.uint32)
    fit = np.zeros(N, np.float32)
    raw = np.zeros(N, np.float32)
    lw = np.zeros(N, np.uint32)
    dec = C.c_uint64()
    game_offsets = (3 * np.concatenate([[0], np.cumsum(np.diff(offsets.astype(np.uint64)))[:-1]])).astype(np.uint64)
    traces = np.zeros(int(3 * offsets[-1]), GAME_TRACE_DTYPE) if trace else None
    st = lib().oro_evaluate(C.cast(ptrs, C.c_void_p), N, enc_hidden, A_genome, A_runtime, ptr(cat), cat.shape[0],
                            ptr(offsets), ptr(ids), mode, threads, ptr(fit), ptr(raw), ptr(lw), C.byref(dec),
                            ptr(game_offsets), ptr(traces))
    return dict(status=st, fitness=fit, raw=raw, local_w=lw, decisions=dec.value, traces=traces,
                game_offsets=game_offsets)


def replay_game(genome, A, catalog, local, slot, guesses, enc_hidden=0):
    g = np.ascontiguousarray(genome, np.float32)
    cat = np.ascontiguousarray(catalog, np.uint8)
    loc = np.ascontiguousarray(local, np.uint32)
    gs = np.ascontiguousarray(guesses, np.uint32)
    n = len(gs)
    pol = np.zeros((max(n, 1), 64), np.float64)
    best = np.zeros(max(n, 1), np.float64)
    second = np.zeros(max(n, 1), np.float64)
    arg = np.zeros(max(n, 1), np.uint32)
    chosen = np.zeros(max(n, 1), np.float64)
    st = lib().oro_replay_game(ptr(g), enc_hidden, A, ptr(cat), ptr(loc), loc.size, slot, ptr(gs), n, ptr(pol),
                               ptr(best), ptr(second), ptr(arg), ptr(chosen))
    return dict(status=st, policy=pol[:n], best=best[:n], second=second[:n], argmax=arg[:n], chosen=chosen[:n])


def summary(fitness):
    f = np.ascontiguousarray(fitness, np.float32)
    b = C.c_float()
    m = C.c_float()
    i = C.c_uint64()
    lib().oro_summary(ptr(f), f.size, C.byref(b), C.byref(m), C.byref(i))
    return b.value, m.value, i.value


def rank_weights(exponent, n):
    w = np.zeros(n + 1, np.float32)
    lib().oro_rank_weights(exponent, n, ptr(w))
    return w


NO_INDEX = 0xFFFFFFFFFFFFFFFF


def select_one(fitness, has, candidates, wtable, tie_seed, u, excluded=NO_INDEX):
    """One roulette draw over `candidates` (cell indices, in list order).  Returns the chosen cell or None."""
    f = np.ascontiguousarray(fitness, np.float32)
    h = np.ascontiguousarray(has, np.uint8)
    c = np.ascontiguousarray(candidates, np.uint64)
    w = np.ascontiguousarray(wtable, np.float32)
    out = C.c_uint64()
    ok = lib().oro_select_one(ptr(f), ptr(h), ptr(c), c.size, ptr(w), tie_seed, u, excluded, C.byref(out))
    return out.value if ok else None


def plan(fitness, has, cur_size, child_count, cols, gen_index, radius, wtable, generation_seed):
    f = np.ascontiguousarray(fitness, np.float32)
    h = np.ascontiguousarray(has, np.uint8)
    w = np.ascontiguousarray(wtable, np.float32)
    pairs = np.zeros((child_count, 2), np.uint32)
    st = lib().oro_plan(ptr(f), ptr(h), cur_size, child_count, cols, gen_index, radius, ptr(w), generation_seed,
                        ptr(pairs))
    return st, pairs


def plan_brood(fitness, has, cur_size, brood, cols, gen_index, radius, wtable, generation_seed):
    """Parent pairs of a screened step's brood: brood x cur_size children, child c of cell c % cur_size."""
    f = np.ascontiguousarray(fitness, np.float32)
    h = np.ascontiguousarray(has, np.uint8)
    w = np.ascontiguousarray(wtable, np.float32)
    pairs = np.zeros((brood * cur_size, 2), np.uint32)
    st = lib().oro_plan_brood(ptr(f), ptr(h), cur_size, brood, cols, gen_index, radius, ptr(w), generation_seed,
                              ptr(pairs))
    return st, pairs


def xover_decisions(A_parent, seed, gen, child, t1, t2, t3, enc_hidden=0):
    rec = np.zeros(int(lib().oro_xover_record_words()), np.uint32)
    lib().oro_xover_decisions(enc_hidden, A_parent, seed, gen, child, t1, t2, t3, ptr(rec))
    return rec


def breed_child(p1, p2, A_parent, A_child, seed, gen, child, t1, t2, t3, p, sigma, p_scale, enc_hidden=0):
    n_child = policy_param_count(enc_hidden) + 38 * A_child
    a = np.ascontiguousarray(p1, np.float32)
    b = np.ascontiguousarray(p2, np.float32)
    out = np.zeros(n_child, np.float32)
    rec = np.zeros(int(lib().oro_xover_record_words()), np.uint32)
    lib().oro_breed_child(enc_hidden, ptr(a), ptr(b), ptr(out), A_parent, A_child, seed, gen, child, t1, t2, t3, p,
                          sigma, p_scale, ptr(rec))
    return out, rec


def bootstrap_genome(A, seed, gen, organism, layer_std, tail_std, enc_hidden=0):
    n = policy_param_count(enc_hidden) + 38 * A
    out = np.zeros(n, np.float32)
    ls = np.ascontiguousarray(layer_std, np.float32)
    lib().oro_bootstrap_genome(enc_hidden, ptr(out), A, seed, gen, organism, ptr(ls), tail_std)
    return out


def layer_stds(variance_scale=2.0, gain=1.0, fan_ins=(145, 321, 256, 128)):
    return np.array([np.float32(gain) * np.sqrt(np.float32(variance_scale) / np.float32(f)) for f in fan_ins],
                    np.float32)


def encode_turn(genome, letters, tiles, enc_hidden=0):
    g = np.ascontiguousarray(genome, np.float32)
    l = np.ascontiguousarray(letters, np.uint8)
    t = np.ascontiguousarray(tiles, np.uint8)
    o32 = np.zeros(64, np.float32)
    o64 = np.zeros(64, np.float64)
    lib().oro_encode_turn(ptr(g), enc_hidden, ptr(l), ptr(t), ptr(o32), ptr(o64))
    return o32, o64


def fixed_features(word_letters):
    w = np.ascontiguousarray(word_letters, np.uint8)
    out = np.zeros(26, np.float32)
    lib().oro_fixed_features(ptr(w), ptr(out))
    return out


def score_action(policy, fixed, tail):
    return lib().oro_score_action(ptr(np.ascontiguousarray(policy, np.float32)),
                                  ptr(np.ascontiguousarray(fixed, np.float32)),
                                  ptr(np.ascontiguousarray(tail, np.float32)))


def slab_after_bootstrap(slot_count, n):
    """Free stack / count / slot states after creation and a bootstrap of n organisms."""
    stack = np.array([slot_count - 1 - i for i in range(slot_count)], np.uint32)
    state = np.zeros(slot_count, np.uint32)
    state[:n] = 1
    return stack, np.uint32(slot_count - n), state, np.arange(n, dtype=np.uint32)


def assemble_slots(slot_count, stack, free_count, state, cur_slots, pairs, spill_capacity=0):
    stack = stack.copy()
    state = state.copy()
    cur = np.ascontiguousarray(cur_slots, np.uint32).copy()
    pr = np.ascontiguousarray(pairs, np.uint32)
    fc = np.array([free_count], np.uint32)
    child = np.zeros(len(pr), np.uint32)
    nb = np.zeros(1, np.uint32)
    sp = np.zeros(1, np.uint32)
    st = lib().oro_assemble_slots(slot_count, ptr(stack), ptr(fc), ptr(state), ptr(cur), cur.size, ptr(pr), len(pr),
                                  ptr(child), ptr(nb), spill_capacity, ptr(sp))
    return dict(status=st, stack=stack, free_count=int(fc[0]), state=state, parent_slots=cur, child_slots=child,
                batches=int(nb[0]), spilled=int(sp[0]))


def hint_group(word_letters):
    w = np.ascontiguousarray(word_letters, np.uint8)
    tc = np.zeros(3, np.uint32)
    letters = np.zeros((3, 5, 5), np.uint8)
    tiles_ = np.zeros((3, 5, 5), np.uint8)
    ok = lib().oro_hint_group(ptr(w), ptr(tc), ptr(letters), ptr(tiles_))
    return ok, tc, letters, tiles_


def median_tail_norm(tails):
    t = np.ascontiguousarray(tails, np.float32)
    return lib().oro_median_tail_norm(ptr(t), t.size // 38)


def inject(genome, parent_a, catalog, first_word, count, mode=0, enc_hidden=0):
    """Returns (status, genome with rows [parent_a, parent_a+count) injected)."""
    g = np.ascontiguousarray(genome, np.float32)
    out = g.copy()
    cat = np.ascontiguousarray(catalog, np.uint8)
    st = lib().oro_inject(ptr(g), ptr(out), enc_hidden, parent_a, ptr(cat), first_word, count, mode)
    return st, out


def repack_slots(new_slot_count, cur_slots, pairs):
    cur = np.ascontiguousarray(cur_slots, np.uint32).copy()
    pr = np.ascontiguousarray(pairs, np.uint32)
    state = np.zeros(new_slot_count, np.uint32)
    stack = np.zeros(new_slot_count, np.uint32)
    fc = np.zeros(1, np.uint32)
    moves = np.zeros(len(cur), np.uint32)
    k = np.zeros(1, np.uint32)
    st = lib().oro_repack_slots(new_slot_count, ptr(cur), cur.size, ptr(pr), len(pr), ptr(state), ptr(stack), ptr(fc),
                                ptr(moves), ptr(k))
    return dict(status=st, parent_slots=cur, state=state, stack=stack, free_count=int(fc[0]),
                moves=moves[:int(k[0])])
