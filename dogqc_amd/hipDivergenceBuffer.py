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
        if self.bufferType is BufferType.REG:
            self._consumeReg(ctxt, pipe, X, live, cnt, thr, outer)
        else:
            self._consumeSmem(ctxt, pipe, X, live, cnt, thr, outer)

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
