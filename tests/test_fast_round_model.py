"""A CPU model of the FAST collide engine's round protocol (engine.hip:
engine_dfs, flush_deferred, dump_deferred and the round loop of
engine_collide), run under many seeded interleavings of its wavefronts.

The GPU tests check the engine's outputs; whether a round boundary falls
mid-stack, a deferred sub-buffer fills, or the idle stop fires depends on
timing there.  Here the same protocol runs step by step on the CPU with a
random wavefront chosen for each step, over a synthetic traversal tree,
with capacities small enough that every boundary case happens: stack
hand-over (both the LDS ring part and the part spilled to device memory),
unconsumed input, deferred sub-buffers padded with kNoTask once full, the
dump regions, the idle stop, uncertain tasks whose children cross into the
next round, and the inline triangle tests.  Each run must test every task
of the tree exactly once and every leaf pair exactly once: the same
multiset as a sequential depth-first walk.

The model mirrors the kernel's step: pop min(sp, 64) tasks, read the LDS
ring back from the spilled part first when the pop reaches below it, push
two children per overlapping internal pair at sp + 2 * rank after spilling
the deepest ring entries (64 at a time, down to half the limit) when the
push would overflow the ring, queue leaf pairs and uncertain tasks.  The
ring is an array indexed by position modulo its size, as in LDS, so an
overwrite of a live entry would lose or duplicate tasks and fail the check.
"""
import functools
import random
from collections import Counter

import pytest

WAVE = 64
NO_TASK = -1
SUB = 8          # sub-buffers per deferred buffer (one per XCD)
UNC_Q = 128      # uncertain queue per wavefront (LDS)
LEAF_Q_INLINE = 2 * WAVE
LEAF_Q = 256


class Tree:
    """Synthetic BVTT: task (q, node) with children (q, 2 node + 1) and
    (q, 2 node + 2).  Outcomes come from a seeded hash: the FP32 test says
    none / push / leaf / uncertain; an uncertain task's FP64 test says none,
    push or leaf."""

    def __init__(self, seed, depth, p_push, p_unc):
        self.seed, self.depth, self.p_push, self.p_unc = seed, depth, p_push, p_unc

    def _u(self, task, salt):
        return random.Random(hash((self.seed, salt) + task)).random()

    def _level(self, node):
        return (node + 1).bit_length() - 1

    def exact(self, task):
        """The FP64 outcome: 'none', 'push' or 'leaf'."""
        q, node = task
        lvl = self._level(node)
        u = self._u(task, 1)
        if lvl >= self.depth:
            return "leaf" if u < 0.7 else "none"
        p = self.p_push if lvl > 1 else 1.0
        return "push" if u < p else "none"

    def fp32(self, task):
        """The FP32 cull: the exact outcome, or 'unsure'."""
        if self._level(task[1]) > 0 and self._u(task, 2) < self.p_unc:
            return "unsure"
        return self.exact(task)

    @staticmethod
    def children(task):
        q, node = task
        return (q, 2 * node + 1), (q, 2 * node + 2)

    def sequential(self, nq):
        """Every task tested and every leaf pair, by a sequential walk."""
        tested, leaves = Counter(), Counter()
        for q in range(nq):
            st = [(q, 0)]
            while st:
                t = st.pop()
                tested[t] += 1
                r = self.exact(t)
                if r == "push":
                    c0, c1 = self.children(t)
                    st += [c1, c0]
                elif r == "leaf":
                    leaves[t] += 1
        return tested, leaves


class Engine:
    """The round protocol with its capacities as parameters."""

    def __init__(self, tree, rng, nwaves, inline, stack_top, stack_cap, defer_leaf,
                 defer_unc, idle_stop, stop_poll, bug=None):
        self.tree, self.rng, self.bug = tree, rng, bug
        self.nwaves, self.inline = nwaves, inline
        self.top_lim, self.stack_cap = stack_top, stack_cap
        self.top_size = 512
        self.lq_cap = LEAF_Q_INLINE if inline else LEAF_Q
        self.dsub = [max(LEAF_Q, defer_leaf // SUB), max(UNC_Q, defer_unc // SUB)]
        self.idle_stop, self.stop_poll = idle_stop, stop_poll
        self.tested, self.leaves = Counter(), Counter()
        self.rounds = 0
        self.events = Counter()

    # -- device-side helpers -------------------------------------------
    def flush_deferred(self, w, x, q, n, k):
        s = self.resv[w][x]
        self.resv[w][x] += k
        cap = self.dsub[w]
        dst = self.defer[w][x]
        if s + k > cap:
            for i in range(s, cap):
                dst[i] = NO_TASK
            self.events["subbuffer_full"] += 1
            return False, n
        for i in range(k):
            dst[s + i] = q[n - k + i]
        return True, n - k

    def dump_deferred(self, w, q, n):
        if n == 0:
            return
        self.dump[w].extend(q[:n])
        assert len(self.dump[w]) <= self.nwaves * [LEAF_Q, UNC_Q][w], "dump region overflow"
        self.events["dump"] += 1

    def leaf_test(self, task):
        self.leaves[task] += 1

    # -- one wavefront as a generator: one yield per step ---------------
    def wave(self, wid):
        tree = self.tree
        xcd = wid % SUB
        ring = [None] * self.top_size
        gstack = [None] * self.stack_cap
        sp = base = 0
        lq, uq = [None] * self.lq_cap, [None] * UNC_Q
        nl = nu = 0
        input_left, stopped, step = True, False, 0
        lim = self.top_lim

        def slot_get(pos):
            assert base <= pos < sp, "read outside the live ring part"
            v = ring[pos % self.top_size]
            assert v is not None
            return v

        while True:
            step += 1
            if step % self.stop_poll == 0 and self.stop:
                stopped = True
                break
            if self.inline and nl >= WAVE:
                nl -= WAVE
                for t in lq[nl:nl + WAVE]:
                    self.leaf_test(t)
                self.events["inline_leaf_step"] += 1
                yield
                continue
            full = False
            if nl >= self.lq_cap // 2:
                ok, nl = self.flush_deferred(0, xcd, lq, nl, self.lq_cap // 2)
                full = not ok
            if not full and nu >= UNC_Q // 2:
                ok, nu = self.flush_deferred(1, xcd, uq, nu, UNC_Q // 2)
                full = not ok
            if full:
                self.stop = True
                stopped = True
                break
            if sp < WAVE and input_left:
                want = WAVE - sp
                start = self.head
                self.head += want
                got = max(0, min(want, self.n_in - start))
                if got < want:
                    input_left = False
                for i in range(got):
                    if self.inline:
                        assert ring[(sp + i) % self.top_size] is None
                        ring[(sp + i) % self.top_size] = self.front_in[start + i]
                    else:
                        gstack[sp + i] = self.front_in[start + i]
                sp += got
            if sp == 0:
                break
            n = min(sp, WAVE)
            if self.inline and sp - n < base:
                half = lim // 2
                r = min(base, half)
                for i in range(r):
                    assert ring[(base - r + i) % self.top_size] is None
                    ring[(base - r + i) % self.top_size] = gstack[base - r + i]
                base -= r
                self.events["readback"] += 1
                assert sp - base <= self.top_size
            tasks = [slot_get(sp - n + i) if self.inline else gstack[sp - n + i]
                     for i in range(n)]
            sp -= n
            if self.inline:
                for i in range(n):
                    ring[(sp + i) % self.top_size] = None  # popped
            res = []
            for t in tasks:
                self.tested[t] += 1
                res.append(tree.fp32(t))
            pushes = [t for t, r in zip(tasks, res) if r == "push"]
            npush = 2 * len(pushes)
            if sp + npush > self.stack_cap:
                raise AssertionError("stack overflow")
            if self.inline and sp + npush - base > lim:
                wcount = sp + npush - base - lim // 2
                wcount = (wcount + WAVE - 1) // WAVE * WAVE
                wcount = min(wcount, sp - base)
                for i in range(wcount):
                    gstack[base + i] = ring[(base + i) % self.top_size]
                    ring[(base + i) % self.top_size] = None
                base += wcount
                self.events["spill"] += 1
            if self.inline:
                assert sp + npush - base <= lim <= self.top_size
            for k, t in enumerate(pushes):
                c0, c1 = tree.children(t)
                o = sp + 2 * k
                if self.inline:
                    for p, c in ((o, c0), (o + 1, c1)):
                        assert ring[p % self.top_size] is None, "live ring entry overwritten"
                        ring[p % self.top_size] = c
                else:
                    gstack[o], gstack[o + 1] = c0, c1
            sp += npush
            for t, r in zip(tasks, res):
                if r == "leaf":
                    lq[nl] = t
                    nl += 1
                elif r == "unsure":
                    uq[nu] = t
                    nu += 1
            assert nl <= self.lq_cap and nu <= UNC_Q
            yield
        if stopped:
            for i in range(sp):
                if self.bug == "handover_ring_only" and i < base:
                    continue
                if self.inline and i >= base:
                    self.front_out.append(ring[i % self.top_size])
                else:
                    self.front_out.append(gstack[i])
            if sp:
                self.events["stack_handover"] += 1
                if self.inline and base > 0:
                    self.events["handover_with_spill"] += 1
            self.dump_deferred(0, lq, nl)
            self.dump_deferred(1, uq, nu)
        else:
            if nl > 0:
                ok, nl2 = self.flush_deferred(0, xcd, lq, nl, nl)
                if not ok:
                    self.dump_deferred(0, lq, nl)
            if nu > 0:
                ok, nu2 = self.flush_deferred(1, xcd, uq, nu, nu)
                if not ok:
                    self.dump_deferred(1, uq, nu)
            self.idle += 1
            if self.idle_limit and self.idle == self.idle_limit:
                self.stop = True
                self.events["idle_stop"] += 1

    # -- host side -----------------------------------------------------
    def run_round(self, front):
        self.rounds += 1
        assert self.rounds < 10000, "no progress"
        self.front_in, self.n_in = front, len(front)
        self.head = 0
        self.front_out = []
        self.resv = [[0] * SUB for _ in range(2)]
        self.defer = [[[None] * self.dsub[w] for _ in range(SUB)] for w in range(2)]
        self.dump = [[], []]
        self.stop = False
        self.idle = 0
        big = len(front) >= WAVE * self.nwaves
        self.idle_limit = max(1, int(self.idle_stop * self.nwaves)) if big else 0
        gens = {w: self.wave(w) for w in range(self.nwaves)}
        live = list(gens)
        while live:
            w = self.rng.choice(live)
            try:
                next(gens[w])
            except StopIteration:
                live.remove(w)
        nxt = list(self.front_out)
        nxt += front[min(self.head, len(front)):]  # unconsumed input
        # what each deferred buffer holds: the filled part of every
        # sub-buffer (padding included) and the dump region
        held = []
        for w in range(2):
            seq = []
            for x in range(SUB):
                seq += self.defer[w][x][:min(self.resv[w][x], self.dsub[w])]
            seq += self.dump[w]
            assert None not in seq, "unwritten deferred entry"
            held.append([t for t in seq if t != NO_TASK])
            if len(seq) != len(held[w]):
                self.events["padding_skipped"] += 1
        for t in held[1]:  # engine_uncertain: the FP64 test decides
            r = self.tree.exact(t)
            if r == "push":
                nxt += list(self.tree.children(t))
                self.events["uncertain_children"] += 1
            elif r == "leaf":
                self.leaf_test(t)
        for t in held[0]:  # engine_leaves
            self.leaf_test(t)
        return nxt

    def expand(self, front, target, levels):
        """Breadth-first levels while the front is small (engine_expand)."""
        for _ in range(levels):
            if not front or len(front) >= target:
                break
            out = []
            for t in front:
                self.tested[t] += 1
                r = self.tree.exact(t)  # FP32 with FP64 when unsure: the exact outcome
                if r == "push":
                    out += list(self.tree.children(t))
                elif r == "leaf":
                    self.leaf_test(t)
            front = out
        return front

    def run(self, nq, target, levels):
        front = self.expand([(q, 0) for q in range(nq)], target, levels)
        while front:
            front = self.run_round(front)
            if front and len(front) < WAVE * self.nwaves:
                front = self.expand(front, target, levels)


DEPTH, P_PUSH, P_UNC, NQ = 11, 0.8, 0.05, 24
CASES = [
    # inline, stack_top, defer_leaf, defer_unc, idle_stop, stop_poll, wavefronts
    # (more than 8 wavefronts share sub-buffers: a flush that only partly
    # fits pads the rest with kNoTask)
    (True, 128, 1 << 12, 64, 0.5, 16, 4),
    (True, 128, 1 << 12, 64, 0.5, 1, 16),
    (True, 256, 1 << 12, 1024, 0.25, 4, 4),
    (True, 512, 1 << 12, 64, 0.0, 1, 12),
    (False, 512, 256, 64, 0.5, 1, 16),
    (False, 512, 2048, 2048, 0.5, 16, 4),
]


@functools.lru_cache(maxsize=None)
def _run(case, seed):
    inline, top, dleaf, dunc, idle, poll, nw = CASES[case]
    tree = Tree(seed=seed * 7 + case, depth=DEPTH, p_push=P_PUSH, p_unc=P_UNC)
    eng = Engine(tree, random.Random(1000 * case + seed), nwaves=nw, inline=inline,
                 stack_top=top, stack_cap=4096, defer_leaf=dleaf, defer_unc=dunc,
                 idle_stop=idle, stop_poll=poll)
    eng.run(NQ, target=64, levels=3)
    return tree, eng


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_round_protocol_tests_every_task_once(case, seed):
    tree, eng = _run(case, seed)
    want_tested, want_leaves = tree.sequential(NQ)
    assert eng.tested == want_tested
    assert eng.leaves == want_leaves
    assert sum(want_leaves.values()) > 100


def test_round_protocol_reaches_every_boundary_case():
    """Across the cases and seeds above, every boundary the protocol has
    happens at least once (so the equality checks cover them)."""
    seen = Counter()
    for case in range(len(CASES)):
        for seed in range(4):
            _, eng = _run(case, seed)
            seen.update(eng.events)
            seen["rounds>1"] += eng.rounds > 1
    for ev in ("subbuffer_full", "dump", "stack_handover", "handover_with_spill", "spill",
               "readback", "idle_stop", "uncertain_children", "inline_leaf_step",
               "padding_skipped", "rounds>1"):
        assert seen[ev] > 0, ev


def test_model_detects_a_planted_handover_mistake():
    """The checks catch a protocol mistake: a stopped wavefront that hands
    over only the LDS part of its stack loses the spilled tasks."""
    lost = 0
    for seed in range(4):
        tree = Tree(seed=seed, depth=DEPTH, p_push=P_PUSH, p_unc=P_UNC)
        want_tested, want_leaves = tree.sequential(NQ)
        eng = Engine(tree, random.Random(seed), nwaves=4, inline=True, stack_top=128,
                     stack_cap=4096, defer_leaf=1 << 12, defer_unc=64, idle_stop=0.5,
                     stop_poll=1, bug="handover_ring_only")
        eng.run(NQ, target=64, levels=3)
        lost += eng.tested != want_tested
    assert lost > 0
