"""CPU model of the SOR launch protocol of csrc/hip/sor_rb.h (sweeping workgroups and the deciding
workgroup of each launch), with every iterate plane holding the number of sweeps applied to it and
a synthetic sequence of per-sweep L1 updates.  It checks the properties the GPU path relies on:

- FP64 (three planes, every sweep tested, rollback): a solve ends with the iterate of the first
  sweep whose update is below tol, and with k equal to that sweep (src/poisson.c:272-278), or
  fails at exactly itmax sweeps; launches queued after the end change nothing;
- FP32 inner solves (two planes, the last sweep of every (test_mask + 1)-th launch tested): the
  final plane holds the iterate of the reported sweep count, and the solve ends with the first
  tested launch whose last update is below tol (or fails at exactly itmax sweeps);
- replayed captured chunks, whose launch index restarts at 0 every M launches, give the same
  result as one continuous sequence when M is a multiple of 6 (FP64) or 12 (FP32, test_mask 3):
  the slot (mod 2), plane (mod 3 or 2) and test (mod 4) rotations continue across replays.  With
  a chunk length that breaks the test cadence the results differ, so the constraint is needed."""
import random

RUN, CONV, FAIL, RB = "running", "converged", "failed", "rollback"


def launch_indices(chunks):
    """Launch indices of a solve: chunks is None (eager: 0, 1, 2, ...) or a chunk length M (each
    replay of a captured chunk runs L = 0 .. M-1)."""
    L = 0
    while True:
        yield L
        L += 1
        if chunks is not None and L == chunks:
            L = 0


def solve(S, fp64, e_seq, tol, itmax, test_mask=0, chunks=None, extra=2, cap=20000):
    """Runs the protocol until the slot a launch writes is terminal, plus `extra` launches queued
    after the end; returns (status, k, sweeps held by the final plane)."""
    NB = 3 if fp64 else 2
    planes = [0] + [None] * (NB - 1)           # sweeps applied; None: never written
    slots = [None, dict(k=0, parity=0, status=RUN, s_done=0, rb_src=0, rb_n=0)]
    partial = [None, None]                     # per-sweep updates left by the launch of that parity
    gen = launch_indices(chunks)
    after_end = None
    for _ in range(cap):
        L = next(gen)
        prev = slots[(L + 1) & 1]
        # sweepers: sweep on from plane L % NB into (L + 1) % NB, or redo a rollback
        tested = fp64 or (L & test_mask) == test_mask
        s_run, src, dst, rollback = 0, None, None, False
        if prev["status"] == RUN:
            s_run = min(S, itmax - (prev["k"] + prev["s_done"]))
            src, dst = L % NB, (L + 1) % NB
        elif prev["status"] == RB:
            s_run, src, dst, rollback = prev["rb_n"], prev["rb_src"], prev["parity"], True
        new_planes = list(planes)
        mine = None
        if s_run > 0:
            x = planes[src]
            assert x is not None, ("a launch read a plane nothing wrote", L, src)
            es = []
            for _ in range(s_run):
                x += 1
                es.append(update(e_seq, x))
            new_planes[dst] = x
            if not rollback and tested:
                mine = es if fp64 else [es[-1]]
        # decider of launch L: decides on launch L-1 from the partials it left
        P = dict(prev)
        if P["status"] != RUN:
            nxt = P
            if P["status"] == RB:
                nxt["status"] = CONV
            nxt["s_done"] = 0
        else:
            sd = P["s_done"]
            nxt = dict(P, parity=L % NB, status=RUN)
            tested_prev = fp64 or ((L - 1) & test_mask) == test_mask
            if sd > 0 and not tested_prev:
                nxt["k"] = P["k"] + sd
                if nxt["k"] >= itmax:
                    nxt["status"] = FAIL
            elif sd > 0:
                etot = partial[(L + 1) & 1]
                s_conv = -1
                for i, e in enumerate(etot):
                    s = i if fp64 else sd - 1
                    if s_conv < 0 and e < tol:
                        s_conv = s
                if 0 <= s_conv < sd - 1:
                    nxt.update(status=RB, k=P["k"] + s_conv + 1, rb_src=(L + NB - 1) % NB, rb_n=s_conv + 1)
                else:
                    nxt["k"] = P["k"] + sd
                    if s_conv >= 0:
                        nxt["status"] = CONV
                    elif nxt["k"] >= itmax:
                        nxt["status"] = FAIL
            elif itmax <= P["k"]:
                nxt["status"] = FAIL
            budget = itmax - (P["k"] + sd)
            nxt["s_done"] = min(S, max(budget, 0)) if nxt["status"] == RUN else 0
        slots[L & 1] = nxt
        planes = new_planes
        if mine is not None:  # untested and rollback launches leave the buffer as it was
            partial[L & 1] = mine
        if after_end is None and nxt["status"] in (CONV, FAIL):
            after_end = 0
        elif after_end is not None:
            after_end += 1
        if after_end is not None and after_end >= extra:
            fin = slots[L & 1]
            return fin["status"], fin["k"], planes[fin["parity"]]
    raise AssertionError("the solve did not end")


def update(e_seq, j):
    """The L1 update of sweep j (from 1); 1.0 past the end of the sequence."""
    return e_seq[j - 1] if j - 1 < len(e_seq) else 1.0


def random_case(rng):
    n = rng.randint(1, 70)
    e_seq = [rng.choice([1.0, 0.5, 2.0]) for _ in range(n)] + [0.001] + [rng.random() for _ in range(40)]
    return e_seq, rng.randint(1, 90)


def test_fp64_stops_at_the_first_sweep_below_tol():
    rng = random.Random(1)
    for _ in range(3000):
        S = rng.choice([1, 2, 4, 5, 6, 8])
        e_seq, itmax = random_case(rng)
        tol = 0.01
        first = next((j + 1 for j, e in enumerate(e_seq) if e < tol), None)
        for chunks in (None, 6, 66):
            status, k, held = solve(S, True, e_seq, tol, itmax, chunks=chunks, extra=rng.randint(0, 8))
            if first is not None and first <= itmax:
                assert (status, k, held) == (CONV, first, first), (S, chunks, first, itmax)
            else:
                assert (status, k, held) == (FAIL, itmax, itmax), (S, chunks, first, itmax)


def test_fp32_tests_every_few_launches():
    rng = random.Random(2)
    for _ in range(3000):
        S = rng.choice([2, 4, 5, 6, 8])
        mask = rng.choice([0, 1, 3])
        e_seq, itmax = random_case(rng)
        tol = 0.01
        status, k, held = solve(S, False, e_seq, tol, itmax, test_mask=mask, extra=rng.randint(0, 8))
        assert held == k
        # the sweep each tested launch ends on: launch L (from 0) ends at min((L + 1) * S, itmax)
        tested_ends = [min((L + 1) * S, itmax) for L in range((itmax + S - 1) // S) if (L & mask) == mask]
        if status == CONV:
            # the first tested launch whose last update is below tol ends the solve
            assert k in tested_ends and update(e_seq, k) < tol
            assert all(update(e_seq, j) >= tol for j in tested_ends if j < k)
        else:
            assert status == FAIL and k == itmax
            assert all(update(e_seq, j) >= tol for j in tested_ends)


def test_replayed_chunks_equal_continuous_launches():
    rng = random.Random(3)
    for _ in range(2000):
        S = rng.choice([4, 5, 6])
        e_seq, itmax = random_case(rng)
        tol = 0.01
        for fp64, mask, chunk in ((True, 0, 6), (True, 0, 12), (False, 3, 12), (False, 3, 72), (False, 1, 6)):
            eager = solve(S, fp64, e_seq, tol, itmax, test_mask=mask)
            replayed = solve(S, fp64, e_seq, tol, itmax, test_mask=mask, chunks=chunk)
            assert eager == replayed, (fp64, mask, chunk, S, itmax)


def test_chunks_that_break_the_test_cadence_change_the_result():
    # the FP32 chunk unit of 12 is needed: with test_mask 3 and chunks of 6 launches the test
    # cadence restarts every replay, and some solves end on another sweep than the eager ones
    rng = random.Random(4)
    differ = 0
    for _ in range(500):
        e_seq, itmax = random_case(rng)
        eager = solve(5, False, e_seq, 0.01, itmax, test_mask=3)
        replayed = solve(5, False, e_seq, 0.01, itmax, test_mask=3, chunks=6)
        differ += eager != replayed
    assert differ > 0
