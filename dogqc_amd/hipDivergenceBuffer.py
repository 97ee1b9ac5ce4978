"""Lane refill for divergent pipelines, designed for 64-lane wavefronts.

Semantics of the reference's LaneRefill / BufferedLoop
(dogqc/cudaDivergenceBuffer.py:11-283): at a plan position, tuples that are
still active are parked in an on-chip buffer, and the operators above that
position run only once at least ``threshold`` x lanes of the (warp) wavefront
can be filled; at the end of the input the buffer drains.  Positions inside
inner loops are skipped (cudaDivergenceBuffer.py:21-24).

Two buffer kinds, selected like the reference's BufferType
(cudaDivergenceBuffer.py:44-46) through ``LaneRefill.bufferType``:

* ``BufferType.SMEM`` (default): every wavefront owns a 128-entry LDS queue
  per live register (attributes already in registers, plus the row indices
  of attributes that are still loaded lazily, so columns that only later
  operators read are fetched after compaction, for dense lanes only).
  Survivors are appended with one 64-bit ballot and an mbcnt prefix; the
  downstream code runs on up to 64 popped entries at a time.

* ``BufferType.REG``: the buffer is one register per live register and lane,
  64 entries per wavefront, no LDS (the reference's register variant,
  cudaDivergenceBuffer.py:73-113, 129-164, 180-209, which its BufferedLoop
  never selects).  When the active lanes plus the buffered entries reach the
  threshold, inactive lanes are refilled from the top of the buffer with one
  ds_bpermute per 32-bit word and the operators above run; otherwise the
  active lanes are pushed onto the buffer with one ds_permute per word, each
  lane sending to a distinct lane (active lanes to the buffer's free
  entries, inactive lanes to the rest), so the permutation is total.

The release threshold is a run-time kernel parameter (1..64 lanes), so the
autotuner can change it without recompiling.
"""
from enum import Enum

from dogqc_amd.translatorBase import UnaryTranslator


class BufferType(Enum):
    SMEM = 1
    REG = 2
    GLOBAL = 3


class LaneRefill(UnaryTranslator):

    threshold = 0.9
    bufferType = BufferType.SMEM

    def __init__(self, parentAlgExpr, child):
        UnaryTranslator.__init__(self, parentAlgExpr, child)
        self.threshold = LaneRefill.threshold
        self.bufferType = LaneRefill.bufferType

    # a column still to be loaded that at least this fraction of the
    # released rows reads later in the pipeline loads right after the
    # release, with the others: the lanes are dense there, and one round
    # trip replaces one per operator that reads a column
    EAGER_AFTER_REFILL = 0.5

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def _eagerAfterRelease(self, pipe):
        from dogqc_amd.hipTranslator import useFractions
        cand = {}
        for id, st in pipe.attrs.items():
            if not st.loaded and st.loader is not None and tuple(st.deps) == (pipe.tid,):
                cand[id] = st.attr
        if not cand:
            return
        fracs = useFractions(self, set(cand), list(cand.values()))
        eager = [cand[id] for id, f in fracs.items() if f >= LaneRefill.EAGER_AFTER_REFILL]
        if eager:
            pipe.need(eager)

    def opId(self):
        return self.algExpr.opId + 100000

    def consume(self, ctxt):
        pipe = ctxt.pipe
        X = self.algExpr.opId
        if ctxt.innerLoopCount > 0 or pipe.innerDepth > 0 or pipe.U != 1:
            pipe.emit("// lane refill at opId %d skipped inside an inner loop" % X)
            self.parent.consume(ctxt)
            return
        pipe.annotations.append("LR%d" % X)
        pipe.loopOps += 1
        live = pipe.liveState(self.algExpr.outRelation.values())
        if pipe.tid not in live:
            live[pipe.tid] = "int64_t"
        cnt = pipe.reg("rcnt%d" % X, "uint32_t", "0", perRow=False)
        default = max(1, min(64, int(64 * self.threshold)))
        pid = pipe.pipeId
        thr = pipe.param("rthr%d" % X, "uint32_t",
                         "std::max<int32_t>(1, std::min<int32_t>(64, t->refill_threshold[%d] > 0 ? "
                         "t->refill_threshold[%d] : %d))" % (pid, pid, default))
        outer = pipe.refillPoints[-1] if pipe.refillPoints else None
        if self.bufferType is BufferType.GLOBAL and self._globalPossible(pipe):
            self._consumeGlobal(ctxt, pipe, X, live)
            return
        if self.bufferType is BufferType.REG:
            self._consumeReg(ctxt, pipe, X, live, cnt, thr, outer)
        else:
            self._consumeSmem(ctxt, pipe, X, live, cnt, thr, outer)

    def _globalPossible(self, pipe):
        """The kernel split needs the first refill point of a pipeline that
        is not launched in phases otherwise and ends in an aggregation or a
        materialisation (a multimap build sink launches in phases itself)."""
        from dogqc_amd.hipTranslator import AggregationTranslator, MaterializeTranslator, ProjectionTranslator
        from dogqc_amd.hashJoins import EquiJoinTranslator
        if pipe.phased or pipe.refillPoints or pipe.partAgg is not None:
            return False
        t = self
        while t.parent is not None:
            p = t.parent
            if isinstance(p, EquiJoinTranslator) and p.leftChild is t:
                return False
            if isinstance(p, (AggregationTranslator, MaterializeTranslator, ProjectionTranslator)):
                return True
            t = p
        return True

    def _consumeGlobal(self, ctxt, pipe, X, live):
        """Lane refill across a kernel boundary (BufferType.GLOBAL): the
        pipeline launches twice.  PHASE 0 runs the operators below this
        point and appends the survivors' live registers to a device queue
        (one atomic per wavefront); PHASE 1 runs over the queue, every lane
        one survivor, so the operators above start with full wavefronts and
        a launch of their own (no LDS queue, no filter in LDS).  The queue
        is sized from the plan's estimate; an overflow is reported through
        split_count and the run repeated with the exact count."""
        prog = ctxt.program
        ht = prog.newHt("queue", X, "gq%d" % X)
        idx = ht["index"]
        est = prog.estimate(self.algExpr)
        prog.hostVars["htcap%d" % idx] = "L.htcap[%d] = 0; L.mment[%d] = 0;" % (idx, idx)
        prog.splitVars["scap%d" % idx] = (
            "L.scap[%d] = t->split_capacity[%d] > 0 ? t->split_capacity[%d] : std::max<int64_t>(4096, %s + %s / 4);"
            % (idx, idx, idx, est, est))
        cap = "L.scap[%d]" % idx
        fcap = pipe.param("gqcap%d" % X, "uint64_t", cap)
        regs = []
        for k, (r, ct) in enumerate(live.items()):
            b = prog.buffer("gq%d_r%d" % (X, k), ct, cap)
            regs.append((r, pipe.param("gq%d_r%d" % (X, k), "%s*" % ct, "B_" + b)))
        pipe.annotations.append("GQ%d" % X)
        # survivors are staged per wavefront in LDS and appended with one
        # atomic per stage flush (one per wavefront and trip on a single
        # counter serialises there: c3 took 10.9 ms)
        from dogqc_amd.hiplang import CType
        NW = pipe.bs // 64
        rowBytes = sum(CType.size.get(ct, 8) for _, ct in live.items())
        SW = max(64, min(128, (2048 // max(1, rowBytes)) // 64 * 64))
        for k, (r, ct) in enumerate(live.items()):
            pipe.staticLds.append("__shared__ %s gqs%d_%d[%d][%d];" % (ct, X, k, NW, SW))
        pipe.staticLds.append("__shared__ uint32_t gqs%d_blk[%d];" % (X, NW))
        pipe.staticLds.append("__shared__ unsigned long long gqs%d_base;" % X)
        gqn = pipe.reg("gqn%d" % X, "uint32_t", "0u", perRow=False)
        counter = "st->split_count[%d]" % idx
        # everything emitted so far is PHASE 0
        pipe.code.lines.insert(0, "    " * (pipe.code.depth - 1) + "if constexpr (PHASE == 0) {")
        pipe.code.depth += 1
        pipe.emit("// lane refill across launches (opId %d): append the survivors" % X)
        pipe.code.open("")
        pipe.emit("const uint64_t qm = dqc::ballot(active[u]);")
        pipe.emit("const uint32_t qn = dqc::popc64(qm);")
        pipe.code.open("if (%s + qn > %du)" % (gqn, SW))
        pipe.emit("dqc::wave_lds_sync();")
        pipe.emit("unsigned long long qb = 0;")
        pipe.emit("if (lane == 0u) qb = atomicAdd((unsigned long long*)&%s, (unsigned long long)%s);"
                  % (counter, gqn))
        pipe.emit("qb = dqc::uniform64(dqc::shfl(qb, 0));")
        pipe.code.open("for (uint32_t i = lane; i < %s; i += 64u)" % gqn)
        pipe.code.open("if (qb + i < %s)" % fcap)
        for k, (r, f) in enumerate(regs):
            pipe.emit("%s[qb + i] = gqs%d_%d[wave][i];" % (f, X, k))
        pipe.code.close()
        pipe.code.close()
        pipe.emit("dqc::wave_lds_sync();")
        pipe.emit("%s = 0u;" % gqn)
        pipe.code.close()
        pipe.code.open("if (active[u])")
        pipe.emit("const uint32_t sp = %s + dqc::lanes_below(qm);" % gqn)
        for k, (r, f) in enumerate(regs):
            pipe.emit("gqs%d_%d[wave][sp] = %s[u];" % (X, k, r))
        pipe.code.close()
        pipe.emit("%s += qn;" % gqn)
        pipe.emit("active[u] = 0;")
        pipe.code.close()
        pipe.code.close()   # end of PHASE 0
        # end of the first launch: the workgroup's staged survivors, one
        # atomic per workgroup
        e = pipe.epilogue
        e.open("if constexpr (PHASE == 0)")
        e.emit("dqc::wave_lds_sync();")
        e.emit("if (lane == 0u) gqs%d_blk[wave] = %s;" % (X, gqn))
        e.emit("__syncthreads();")
        e.open("if (threadIdx.x == 0u)")
        e.emit("uint32_t tot = 0u;")
        e.emit("for (uint32_t w = 0; w < %du; ++w) tot += gqs%d_blk[w];" % (NW, X))
        e.emit("gqs%d_base = tot ? atomicAdd((unsigned long long*)&%s, (unsigned long long)tot) : 0ull;"
               % (X, counter))
        e.close()
        e.emit("__syncthreads();")
        e.emit("uint64_t qwb = gqs%d_base;" % X)
        e.emit("for (uint32_t w = 0; w < wave; ++w) qwb += gqs%d_blk[w];" % X)
        e.open("for (uint32_t i = lane; i < %s; i += 64u)" % gqn)
        e.open("if (qwb + i < %s)" % fcap)
        for k, (r, f) in enumerate(regs):
            e.emit("%s[qwb + i] = gqs%d_%d[wave][i];" % (f, X, k))
        e.close()
        e.close()
        e.close()
        pipe.code.open("else")
        pipe.emit("// the loop index is a queue entry")
        pipe.emit("const uint64_t qi = (uint64_t)%s[u];" % pipe.tid)
        pipe.code.open("if (active[u])")
        for r, f in regs:
            pipe.emit("%s[u] = %s[qi];" % (r, f))
        pipe.code.close()
        pipe.code.close()
        pipe.phased = True
        pipe.phaseBetween = []
        pipe.globalRefill = True
        pipe.phaseRows[1] = (cap, "std::min<uint64_t>(p.st->split_count[%d], p.gqcap%d)" % (idx, X))
        self._eagerAfterRelease(pipe)
        self.parent.consume(ctxt)

    def _flushFlag(self, pipe, X, outer):
        if outer is None:
            pipe.emit("const bool rflush%d = !more;" % X)
        else:
            pipe.emit("const bool rflush%d = rflush%d && %s == 0u;" % (X, outer["id"], outer["cnt"]))

    def _consumeSmem(self, ctxt, pipe, X, live, cnt, thr, outer):
        nw = pipe.bs // 64
        queues = []
        for r, ct in live.items():
            q = "rq%d_%s" % (X, r)
            pipe.staticLds.append("__shared__ %s %s[%d][128];" % (ct, q, nw))
            queues.append((r, q))
        pipe.emit("// lane refill (opId %d)" % X)
        pipe.code.open("")
        pipe.emit("const uint64_t rm = dqc::ballot(active[u]);")
        pipe.emit("const uint32_t rpos = %s + dqc::lanes_below(rm);" % cnt)
        pipe.code.open("if (active[u])")
        for r, q in queues:
            pipe.emit("%s[wave][rpos] = %s[u];" % (q, r))
        pipe.code.close()
        pipe.emit("%s += dqc::popc64(rm);" % cnt)
        # entries written by some lanes are read back by others: order the
        # LDS stores before the pops below
        pipe.emit("dqc::wave_lds_sync();")
        pipe.code.close()
        self._flushFlag(pipe, X, outer)
        point = {"id": X, "cnt": cnt}
        point["mark"] = pipe.code.mark()
        pipe.code.depth += 1
        pipe.refillPoints.append(point)
        pipe.emit("const uint32_t rtake = %s < 64u ? %s : 64u;" % (cnt, cnt))
        pipe.emit("%s -= rtake;" % cnt)
        pipe.emit("active[u] = lane < rtake;")
        pipe.code.open("if (active[u])")
        for r, q in queues:
            pipe.emit("%s[u] = %s[wave][%s + lane];" % (r, q, cnt))
        pipe.code.close()
        # the popped entries are read before later pushes reuse their slots
        pipe.emit("dqc::wave_lds_sync();")
        point["cond"] = "%s >= (rflush%d ? 1u : %s)" % (cnt, X, thr)
        self._eagerAfterRelease(pipe)
        self.parent.consume(ctxt)
        pipe.code.close()
        pipe.emit("active[u] = 0;")

    def _consumeReg(self, ctxt, pipe, X, live, cnt, thr, outer):
        pipe.annotations.append("REG%d" % X)
        bufs = []
        for r, ct in live.items():
            bufs.append((r, pipe.reg("rb%d_%s" % (X, r), ct, "{}", perRow=False)))
        go = pipe.reg("rgo%d" % X, "bool", "false", perRow=False)
        pipe.emit("// lane refill, register buffer (opId %d)" % X)
        self._flushFlag(pipe, X, outer)
        pipe.code.open("")
        pipe.emit("const uint64_t rm = dqc::ballot(active[u]);")
        pipe.emit("const uint32_t rna = dqc::popc64(rm);")
        pipe.emit("%s = %s + rna >= %s || (rflush%d && %s + rna > 0u);" % (go, cnt, thr, X, cnt))
        # below the threshold: park the active lanes at entries cnt.. of the
        # buffer (cnt + rna < thr <= 64, so they fit)
        pipe.code.open("if (!%s)" % go)
        pipe.emit("const uint32_t lb = dqc::lanes_below(rm);")
        pipe.emit("const uint32_t rdst = (%s + (active[u] ? lb : rna + (lane - lb))) & 63u;" % cnt)
        pipe.emit("const bool rin = lane >= %s && lane < %s + rna;" % (cnt, cnt))
        for r, b in bufs:
            pipe.emit("{ const auto rv = dqc::permute_to(%s[u], rdst); if (rin) %s = rv; }" % (r, b))
        pipe.emit("%s += rna;" % cnt)
        pipe.emit("active[u] = 0;")
        pipe.code.close()
        pipe.code.close()
        point = {"id": X, "cnt": cnt}
        point["mark"] = pipe.code.mark()
        pipe.code.depth += 1
        pipe.refillPoints.append(point)
        # fill the inactive lanes from the top of the buffer
        pipe.code.open("")
        pipe.emit("const uint64_t rm = dqc::ballot(active[u]);")
        pipe.emit("const uint32_t rna = dqc::popc64(rm);")
        pipe.emit("const uint32_t rtake = %s < 64u - rna ? %s : 64u - rna;" % (cnt, cnt))
        pipe.emit("const uint32_t rs = lane - dqc::lanes_below(rm);")
        pipe.emit("const bool rdest = !active[u] && rs < rtake;")
        pipe.emit("const uint32_t rsrc = rdest ? %s - rtake + rs : lane;" % cnt)
        for r, b in bufs:
            pipe.emit("{ const auto rv = dqc::shfl(%s, (int)rsrc); if (rdest) %s[u] = rv; }" % (b, r))
        pipe.emit("if (rdest) active[u] = 1;")
        pipe.emit("%s -= rtake;" % cnt)
        pipe.code.close()
        point["cond"] = go
        self._eagerAfterRelease(pipe)
        self.parent.consume(ctxt)
        pipe.emit("active[u] = 0;")
        pipe.emit("%s = %s >= %s || (rflush%d && %s > 0u);" % (go, cnt, thr, X, cnt))
        pipe.code.close()
        pipe.emit("active[u] = 0;")


def finishRefills(pipe):
    """Fill in the loop heads of all refill points of a pipeline: while
    flushing, a loop also runs while any later (downstream) queue still holds
    tuples."""
    points = pipe.refillPoints
    for i, pt in enumerate(points):
        later = points[i + 1:]
        cond = pt["cond"]
        if later:
            pending = " || ".join("%s > 0u" % q["cnt"] for q in later)
            cond = "(%s) || (rflush%d && (%s))" % (cond, pt["id"], pending)
        pipe.code.fill(pt["mark"], "while (%s) {" % cond)
