"""Lane-activity histogram at a plan position (the reference's
LaneActivityProfiler, dogqc/cudaLaneActivityProfile.py:10-43), for 64-lane
wavefronts: 65 counters, counter i = number of wavefront passes through the
position with exactly i active lanes.  Counting is switched on at run time
(``dqc_plan_tuning.profile``), so the same module serves timed and profiled
runs.  The host prints the counters in the reference's ``<pN>`` block format
(``i, count`` lines, cudaLaneActivityProfile.py:22-25).
"""
from dogqc_amd.translatorBase import UnaryTranslator
from dogqc_amd.hiplang import CType


class LaneActivityProfiler(UnaryTranslator):

    def produce(self, ctxt):
        self.child.produce(ctxt)

    def opId(self):
        return self.algExpr.opId + 2000000

    def consume(self, ctxt):
        pipe = ctxt.pipe
        prog = ctxt.program
        P = self.algExpr.opId
        slot = len(prog.profiles)
        prog.profiles.append(P)
        if "profile" not in prog.buffers:
            prog.buffer("profile", CType.UINT64, "%d * DQC_PROFILE_BUCKETS" % 32, "0x0ull")
        pipe.annotations.append("P%d" % P)
        f = pipe.param("prof", "unsigned long long*", "(unsigned long long*)B_profile")
        on = pipe.param("prof_on", "int32_t", "t->profile")
        pipe.emit("// lane activity profile (opId %d)" % P)
        # after a kernel-split lane refill the first launch has no rows
        # here: only the second launch's trips are counted
        cond = on if not getattr(pipe, "globalRefill", False) else "PHASE == 1 && %s" % on
        with pipe.rowLoop():
            pipe.code.open("if (%s)" % cond)
            pipe.emit("const uint64_t pm = dqc::ballot(active[u]);")
            pipe.emit("if (lane == 0) atomicAdd(&%s[%d * DQC_PROFILE_BUCKETS + dqc::popc64(pm)], 1ull);" % (f, slot))
            pipe.code.close()
        self.parent.consume(ctxt)
