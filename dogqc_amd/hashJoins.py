"""Hash equi-join translation (build and probe) for gfx950.

Same decisions as the reference's EquiJoinTranslator (dogqc/hashJoins.py):
the left child builds, the right child probes; ``multimatch`` (cfg, default
True) selects a multimap, otherwise a unique table; INNER/SEMI/ANTI join
types; ``EquiJoinTranslator.usePushDownJoin`` (hashJoins.py:14) selects
push-down parallelism for multimatch probes.  The data structures and the
push-down are designed for 64-lane wavefronts:

  * unique table: 64-bit tags claimed by CAS.  When the key columns pack into
    63 bits the tag is the key itself ("exact" mode): the probe compares one
    word and the build-side key values are taken from the probe-side keys, so
    they are not stored.  Otherwise the tag is a hash and keys are stored and
    compared.  Payload columns are structure-of-arrays and loaded lazily.
  * multimap: bucket directory of (offset, count) pairs, count pass, scan,
    insert pass (hashing.cuh:122-199), entries stored as columns.
  * push-down: the wavefront scans the match-list lengths of its 64 lanes
    and walks the concatenated matches 64 at a time; each lane finds the probe
    tuple owning its match by binary search over the scan in LDS and takes
    that tuple's registers by a lane shuffle.  Every lane does useful work in
    every trip except the last, however skewed the lists (the reference packs
    whole lists into 32 lanes with a leader loop, hashJoins.py:599-681).
  * long lists across wavefronts: push-down balances the lanes of one
    wavefront, but a skewed key (Zipf) can still give one wavefront
    thousands of trips while the rest of the GPU idles.  Probe rows whose
    list exceeds a threshold are queued (their live registers, offset and
    length) instead of walked; after the pass a one-workgroup scan cuts the
    queued lists into chunks of SPLIT_CHUNK matches, and a second launch of
    the same kernel (PHASE 1) gives every lane one chunk and runs the same
    push-down walk and operators over them, spreading a hot key over the
    whole GPU.

Extension over the reference: additional non-equality join conditions are
evaluated for INNER joins too (the reference evaluates them only for
SEMI/ANTI, hashJoins.py:203-204, 264-265).
"""
from collections import OrderedDict

from dogqc_amd.relationalAlgebra import Join
from dogqc_amd.translatorBase import BinaryTranslator
from dogqc_amd.types import Type
from dogqc_amd.hiplang import CType, PACK_BITS


class StoredField(object):
    """One stored build column: column ``name[ix]`` of a unique table's
    structure of arrays, or field k of a multimap's entry records,
    ``name[ix].fk`` (a probe that finds a short match list then reads its
    keys and payloads from one cache line rather than one per column)."""

    def __init__(self, name, k=None):
        self.name, self.k = name, k

    def at(self, ix):
        return "%s[%s]" % (self.name, ix) if self.k is None else "%s[%s].f%d" % (self.name, ix, self.k)


class PackedField(object):
    """The stored column of a packed unique table (one stored column of at
    most 4 bytes, keys of at most 32 bits): the high half of slot ``ix``."""

    def __init__(self, name, ctype):
        self.name, self.ctype = name, ctype

    def at(self, ix):
        return "dqc::packed_payload<%s>(%s[%s])" % (self.ctype, self.name, ix)


class EquiJoinTranslator(BinaryTranslator):

    usePushDownJoin = False
    usePartitionedBuild = False      # unique builds that can be partitioned are translated for it
    # unique tables with 32-bit tags and one stored column of at most 4 bytes
    # keep tag and column in one 64-bit slot (dqc::packed_slot)
    usePackedSlots = True
    # matches per queued chunk: a lane of the second launch takes one chunk,
    # so smaller chunks give that launch more wavefronts (c2 at 16 M probe
    # rows, second launch: 64 -> 76 us, 32 -> 47 us, 16 -> 33 us, 8 -> 31 us)
    SPLIT_CHUNK = 16
    DEFAULT_SPLIT_THRESHOLD = 1024   # list length above which a probe row is queued
    # unique inner joins at most this selective can take a filter of their
    # build keys (tuning join_filter = n KB, up to 4 MB; off by default): in
    # LDS when it fits beside the probe pipeline's other LDS (every workgroup
    # copies it, so it pays with few, large workgroups; the autotuner
    # searches size and grid together), else probed in device memory, where
    # a filter of a few MB stays in L2 while the table it guards does not
    FILTER_SELECTIVITY = 0.25
    # join filter size (KB) a plan uses when its tuning leaves join_filter at
    # 0; plans translated with it carry it in their generated host code
    defaultJoinFilter = 0

    def __init__(self, algExpr, leftChild, rightChild):
        BinaryTranslator.__init__(self, algExpr, leftChild, rightChild)
        self.consumeCall = 0

    def produce(self, ctxt):
        self.leftChild.produce(ctxt)
        self.rightChild.produce(ctxt)

    def consume(self, ctxt):
        self.consumeCall += 1
        if self.consumeCall % 2 == 1:
            self.consumeBuild(ctxt)
        else:
            self.consumeProbe(ctxt)

    # ----------------------------------------------------------- planning
    def _plan(self, ctxt):
        a = self.algExpr
        self.buildKeys = list(a.buildKeyAttributes.values())
        self.probeKeys = list(a.probeKeyAttributes.values())
        leftOut = self.leftChild.algExpr.outRelation
        if a.joinType in (Join.INNER, Join.OUTER):
            above = [att for id, att in a.outRelation.items() if id in leftOut]
        else:
            above = []
        condBuild = list(getattr(a, "conditionBuildAttributes", {}).values())
        self.exact = (all(PACK_BITS.get(b.dataType) for b in self.buildKeys)
                      and sum(PACK_BITS[b.dataType] for b in self.buildKeys) <= 63
                      and all(b.dataType == p.dataType for b, p in zip(self.buildKeys, self.probeKeys)))
        keyIds = set(b.id for b in self.buildKeys)
        stored = OrderedDict()
        if a.multimatch or not self.exact:
            for b in self.buildKeys:
                stored[b.id] = b
        for att in above + condBuild:
            if self.exact and not a.multimatch and att.id in keyIds:
                continue
            stored[att.id] = att
        self.stored = list(stored.values())
        self.keyReconstruct = [b for b in self.buildKeys if b.id not in stored and
                               any(b.id == x.id for x in above + condBuild)]
        # unique tables whose packed keys fit 32 bits use 32-bit tags
        self.tag32 = (self.exact and not a.multimatch and
                      sum(PACK_BITS[b.dataType] for b in self.buildKeys) <= 32)
        self.tagType = "uint32_t" if self.tag32 else "uint64_t"
        # the slot word of a unique table: the tag, or tag and stored column
        # (set by the build, which is translated before the probe)
        self.packed = False
        self.slotType = self.tagType
        # a selective unique join gets an LDS filter of its build keys
        self.filtered = (not a.multimatch and a.joinType == Join.INNER
                         and float(a.selectivity) <= self.FILTER_SELECTIVITY)

    def _preHash(self, pipe):
        """Wavefront string hashes of the probe keys (StringHash.wave), at
        the current uniform position; None when the keys are not hashed."""
        if self.tag32 or self.exact:
            return None
        from dogqc_amd.hipTranslator import hoistStringHashes
        return hoistStringHashes(pipe, self.probeKeys)

    def _hashCode(self, pipe, keys, tagVar, hashVar, tagIsReg=False, pre=None):
        """Emit the tag and hash of a key tuple.  With tagIsReg the tag goes
        into the per-row register ``tagVar`` (for two-phase lookups); `pre`
        holds string key hashes computed by the wavefront
        (hoistStringHashes)."""
        tagDecl = tagVar if tagIsReg else "const %s %s" % (self.tagType, tagVar)
        if self.tag32:
            # packed key + 1; the key with all 32 bits set gets tag 0 (extra slot)
            parts = []
            shift = 0
            for k in keys:
                w = PACK_BITS[k.dataType]
                mask = "0xffffffffu" if w == 32 else "0xffu"
                parts.append("(((uint32_t)(%s) & %s) << %d)" % (pipe.access(k), mask, shift))
                shift += w
            pipe.emit("const uint32_t %s_key = %s;" % (hashVar, " | ".join(parts)))
            pipe.emit("%s = %s_key + 1u;" % (tagDecl, hashVar))
            pipe.emit("const uint64_t %s = dqc::mix64(dqc::TAG_TOP | (uint64_t)%s_key);" % (hashVar, hashVar))
        elif self.exact:
            parts = ["dqc::TAG_TOP"]
            shift = 0
            for k in keys:
                w = PACK_BITS[k.dataType]
                mask = "0xffffffffull" if w == 32 else "0xffull"
                parts.append("(((uint64_t)(%s) & %s) << %d)" % (pipe.access(k), mask, shift))
                shift += w
            pipe.emit("%s = %s;" % (tagDecl, " | ".join(parts)))
            pipe.emit("const uint64_t %s = dqc::mix64(%s);" % (hashVar, tagVar))
        else:
            from dogqc_amd.hipTranslator import _keyBits
            pipe.emit("uint64_t %s = 0;" % hashVar)
            for k in keys:
                pipe.emit("%s = dqc::hash_combine(%s, %s);" % (hashVar, hashVar, _keyBits(k, pipe.access(k), pre)))
            pipe.emit("%s = dqc::TAG_TOP | (%s >> 1);" % (tagDecl, hashVar))

    # ----------------------------------------------------------- build
    def consumeBuild(self, ctxt):
        a = self.algExpr
        pipe = ctxt.pipe
        prog = ctxt.program
        self._plan(ctxt)
        J = a.opId
        pipe.emit("// hash join build (opId %d)" % J)
        pipe.need(self.buildKeys + self.stored)
        est = prog.estimate(self.leftChild.algExpr)
        kind = "multi" if a.multimatch else "unique"
        if a.joinType == Join.OUTER and not a.multimatch:
            raise NotImplementedError("outer joins need a multimatch hash table")
        ht = prog.newHt(kind, J, "jht%d" % J)
        self.ht = ht
        idx = ht["index"]
        prog.hostVars["htcap%d" % idx] = (
            "L.htcap[%d] = t->ht_capacity[%d] > 0 ? t->ht_capacity[%d] : "
            "std::max<int64_t>(1, (int64_t)((double)%s * %r * fac(t->ht_factor[%d])));\n"
            "if (L.htcap[%d] >= (int64_t)0xffffffffll) return DQC_ENOMEM;"
            % (idx, idx, idx, est, float(a.htSizeFactor), idx, idx))
        self.storedBuffers = []
        if kind == "unique":
            prog.hostVars["mment%d" % idx] = "L.mment[%d] = 0;" % idx
            if self._packable(pipe):
                self.packed, self.slotType = True, "uint64_t"
                self.packedType = pipe.attrs[self.stored[0].id].ctype
            # padded by PROBE_WINDOW slots for the probe's window loads (with
            # 32-bit tags the first of them is the extra slot)
            tt = CType.UINT if self.slotType == "uint32_t" else CType.UINT64
            tags = prog.buffer("jht%d_tags" % J, tt, "L.htcap[%d] + 4" % idx, "0x0ull")
            ftags = pipe.param("jht%d" % J, "%s*" % self.slotType, "B_" + tags)
            fcap = pipe.param("jht%d_cap" % J, "uint32_t", "L.htcap[%d]" % idx)
            count = "L.htcap[%d] + 1" % idx
        else:
            prog.hostVars["mment%d" % idx] = (
                "L.mment[%d] = t->mm_entries[%d] > 0 ? t->mm_entries[%d] : std::max<int64_t>(64, %s * 2);\n"
                "if (L.mment[%d] >= (int64_t)0xffffffffll) return DQC_ENOMEM;" % (idx, idx, idx, est, idx))
            dirb = prog.buffer("jht%d_dir" % J, CType.UINT, "2 * L.htcap[%d]" % idx, "0x0ull")
            sums = prog.buffer("jht%d_sums" % J, CType.UINT64, "(L.htcap[%d] + 4095) / 4096 + 1" % idx)
            ht["dirBuffer"], ht["sumsBuffer"] = dirb, sums
            fdir = pipe.param("jht%d_dir" % J, "uint32_t*", "B_" + dirb)
            fcap = pipe.param("jht%d_cap" % J, "uint32_t", "L.htcap[%d]" % idx)
            fecap = pipe.param("jht%d_ecap" % J, "uint32_t", "L.mment[%d]" % idx)
            count = "L.mment[%d]" % idx
        fields = []
        if kind == "unique" and self.packed:
            fields = [PackedField(ftags, self.packedType)]
        elif kind == "unique":
            for k, att in enumerate(self.stored):
                ct = pipe.attrs[att.id].ctype
                b = prog.buffer("jht%d_p%d" % (J, k), ct, count)
                self.storedBuffers.append((b, ct))
                fields.append(StoredField(pipe.param("jht%d_p%d" % (J, k), "%s*" % ct, "B_" + b)))
        else:
            # multimap entries are records of the stored columns, the widest
            # first so that no field needs padding before it
            self.storedTypes = [pipe.attrs[att.id].ctype for att in self.stored]
            rec = "dqc_mm%d" % J
            order = sorted(range(len(self.stored)), key=lambda k: (-CType.size[self.storedTypes[k]], k))
            prog.structDecls.append("struct %s { %s };" % (
                rec, " ".join("%s f%d;" % (self.storedTypes[k], k) for k in order) or "uint32_t unused;"))
            self.entryRecord = rec
            b = prog.buffer("jht%d_e" % J, rec, count)
            fe = pipe.param("jht%d_e" % J, "%s*" % rec, "B_" + b)
            fields = [StoredField(fe, k) for k in range(len(self.stored))]
        fbf = None
        if kind == "unique" and self.filtered:
            # tuning join_filter: n KB, < 0 none, 0 the plan's default
            # (defaultJoinFilter, none unless a plan is translated with one)
            req = "t->join_filter" if not EquiJoinTranslator.defaultJoinFilter else \
                "(t->join_filter != 0 ? t->join_filter : %d)" % int(EquiJoinTranslator.defaultJoinFilter)
            prog.hostVars["bfw%d" % idx] = (
                "{ const int32_t req = " + req + "; uint32_t w = 0u;\n"
                "  if (req > 0) w = std::min<uint32_t>(1u << 20, pow2floor((uint64_t)req * 256u));\n"
                "  L.bfw[%d] = w; }" % idx)
            bf = prog.buffer("jbf%d" % J, CType.UINT, "std::max<uint32_t>(1u, L.bfw[%d])" % idx, "0x0ull")
            fbf = pipe.param("jbf%d" % J, "uint32_t*", "B_" + bf)
            fbfw = pipe.param("jbfw%d" % J, "uint32_t", "L.bfw[%d]" % idx)
        if kind == "unique":
            tt = self.tagType
            ins = {"jtag": pipe.reg("jtag%d" % J, tt), "jc": pipe.reg("jc%d" % J, "uint32_t"),
                   "jprev": pipe.reg("jprev%d" % J, self.slotType), "ftags": ftags, "fcap": fcap, "fields": fields,
                   "bloom": (fbf, fbfw) if fbf is not None else None}
            if self._buildPartitionable(pipe):
                self._partitionedBuild(ctxt, pipe, prog, ins, est)
            else:
                self._directInsert(pipe, ins, "active[u]")
            return
        # multimap: count pass (PHASE 0) and insert pass (PHASE 1) of one kernel
        pipe.phased = True
        pipe.phasedHt = ht
        jb = pipe.reg("jb%d" % J, "uint32_t")
        if pipe.loopOps == 0 and pipe.innerDepth == 0:
            self._prog = prog
            self._multiBuildStaged(pipe, J, idx, fdir, fcap, fecap, fields, jb)
            return
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            self._hashCode(pipe, self.buildKeys, "jtag%d" % J, "jh%d" % J)
            pipe.emit("%s = dqc::slot_of(jh%d, %s);" % (jb, J, fcap))
            pipe.code.open("if constexpr (PHASE == 0)")
            pipe.emit("atomicAdd(&%s[2u * %s + 1u], 1u);" % (fdir, jb))
            pipe.code.close()
            pipe.code.open("else")
            pipe.emit("const uint32_t jpos = %s[2u * %s] + atomicAdd(&%s[2u * %s + 1u], 1u);" % (fdir, jb, fdir, jb))
            self._storeEntry(pipe, idx, fecap, fields, "jpos")
            pipe.code.close()
            pipe.code.close()

    def _packable(self, pipe):
        """Tag and stored column fit one 64-bit slot: 32-bit tags, one
        stored column of at most 4 bytes, and a build that is not translated
        for the partitioned insert (which writes tags and columns apart)."""
        if not (EquiJoinTranslator.usePackedSlots and self.tag32 and len(self.stored) == 1):
            return False
        ct = pipe.attrs[self.stored[0].id].ctype
        if ct == CType.STR_TYPE or CType.size.get(ct, 16) > 4:
            return False
        return not self._buildPartitionable(pipe)

    def _probeFn(self):
        return "dqc::ht_probe_packed" if self.packed else "dqc::ht_probe<%s>" % self.tagType

    def _packedWord(self, pipe, jtag):
        return "dqc::packed_slot(%s, dqc::to_word<%s>(%s))" % (jtag, self.packedType, pipe.access(self.stored[0]))

    def _firstCas(self, pipe, ins):
        """Home slot and first CAS of a row whose tag and hash jhJ are set."""
        J = self.algExpr.opId
        jtag, jc, jprev, ftags, fcap, tt = ins["jtag"], ins["jc"], ins["jprev"], ins["ftags"], ins["fcap"], self.tagType
        if self.packed:
            # tag 0: the extra slot at index cap, with tag 1 in the low half
            pipe.emit("%s = %s != 0u ? dqc::slot_of(jh%d, %s) : %s;" % (jc, jtag, J, fcap, fcap))
            pipe.emit("%s = dqc::cas_tag<uint64_t>(&%s[%s], 0ull, %s);" % (jprev, ftags, jc, self._packedWord(pipe, jtag)))
        elif self.tag32:
            # tag 0: the extra slot at index cap (duplicates of that key are dropped)
            pipe.emit("%s = %s != 0u ? dqc::slot_of(jh%d, %s) : %s;" % (jc, jtag, J, fcap, fcap))
            pipe.emit("%s = dqc::cas_tag(&%s[%s], (%s)0, %s != 0u ? %s : 1u);" % (jprev, ftags, jc, tt, jtag, jtag))
        else:
            pipe.emit("%s = dqc::slot_of(jh%d, %s);" % (jc, J, fcap))
            pipe.emit("%s = dqc::cas_tag(&%s[%s], (%s)0, %s);" % (jprev, ftags, jc, tt, jtag))

    def _resolveInsert(self, pipe, ins, cond, values):
        """After the first CAS: collisions, then the payload columns."""
        idx = self.ht["index"]
        jtag, jc, jprev, ftags, fcap = ins["jtag"], ins["jc"], ins["jprev"], ins["ftags"], ins["fcap"]
        with pipe.rowLoop():
            pipe.code.open("if (%s)" % cond)
            if self.packed:
                pipe.emit("const int64_t js = %s == 0ull ? (int64_t)%s : %s == 0u ? (int64_t)-2 : "
                          "dqc::ht_insert_from<uint64_t>(%s, %s, %s, dqc::next_slot(%s, %s), 1u, &st->flags);"
                          % (jprev, jc, jtag, ftags, fcap, self._packedWord(pipe, jtag), jc, fcap))
            elif self.tag32:
                pipe.emit("const int64_t js = %s == 0u ? (int64_t)%s : %s == 0u ? (int64_t)-2 : "
                          "dqc::ht_insert_from<uint32_t>(%s, %s, %s, dqc::next_slot(%s, %s), 1u, &st->flags);"
                          % (jprev, jc, jtag, ftags, fcap, jtag, jc, fcap))
            else:
                pipe.emit("const int64_t js = %s == 0ull ? (int64_t)%s : dqc::ht_insert_from<uint64_t>(%s, %s, %s, "
                          "dqc::next_slot(%s, %s), 1u, &st->flags);" % (jprev, jc, ftags, fcap, jtag, jc, fcap))
            if not self.packed:
                pipe.code.open("if (js >= 0)")
                for f, v in zip(ins["fields"], values):
                    pipe.emit("%s = %s;" % (f.at("js"), v))
                pipe.code.close()
            pipe.emit("%sif (js == -1) { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); "
                      "atomicOr(&st->ht_overflow_mask, %du); }" % ("" if self.packed else "else ", 1 << idx))
            pipe.code.close()

    def _directInsert(self, pipe, ins, cond):
        """First CAS of every row of the thread first, collisions after."""
        J = self.algExpr.opId
        with pipe.rowLoop():
            pipe.code.open("if (%s)" % cond)
            self._hashCode(pipe, self.buildKeys, ins["jtag"], "jh%d" % J, tagIsReg=True)
            self._firstCas(pipe, ins)
            if ins["bloom"] is not None:
                fbf, fbfw = ins["bloom"]
                pipe.emit("if (%s) dqc::bloom_set(%s, %s, jh%d);" % (fbfw, fbf, fbfw, J))
            pipe.code.close()
        self._resolveInsert(pipe, ins, cond, [pipe.access(att) for att in self.stored])

    # partitioned build (tuning part_build = n partitions; off by default:
    # measured slower than the CAS build on c1's 300 k-key dim table, 45 vs
    # 25 us, and even at 3 M keys, 222 vs 213 us; faster on c3's 2 M-row
    # customer build at 10x, 143 vs 250 us): the LDS image of a partition's
    # slot range may take at most this many bytes
    PART_BUILD_SPAN_MAX = 49152

    def _buildPartitionable(self, pipe):
        """A unique build can be launched in two phases when its insert is not
        nested in a loop of the pipeline, the keys are exact (the hash follows
        from the tag) and at most 8 columns are stored, each in at most 8
        bytes (a record word; not strings)."""
        if not EquiJoinTranslator.usePartitionedBuild:
            return False
        if pipe.innerDepth or pipe.loopOps or pipe.phased or pipe.refillPoints:
            return False
        if any(l is None for l in pipe.code.lines):
            return False
        if not all(CType.size.get(pipe.attrs[att.id].ctype, 16) <= 8 for att in self.stored):
            return False
        return self.exact and len(self.stored) <= 8

    def _partitionedBuild(self, ctxt, pipe, prog, ins, est):
        """Unique build without a global CAS per key, two launches of the
        build pipeline plus a spill step.

        Device-scope CAS is what bounds a hash build on MI355X (c1's dim
        build: 300 k inserts, 25 us; with plain stores instead, 9 us).
        PHASE 0 runs the pipeline; a row's home slot decides its partition
        (np contiguous slot ranges of `span` slots), and workgroup w of the
        launch appends (tag, stored columns) records to its own segment of
        that partition through an LDS cursor.  Rows that do not fit their
        segment, and the tag-0 key, are inserted directly (CAS).  PHASE 1
        gives workgroup b partition b: it copies the tags of its slot range
        into LDS (slots taken by direct inserts stay taken), claims each
        record's first free slot from its home slot in that LDS image, and
        writes the tag and stored columns to the claimed slot with plain
        stores -- no other workgroup writes that range.  A record whose probe
        run leaves the range goes to a spill list that a last kernel inserts
        with the ordinary CAS probe; every slot from its home slot to the
        range end is taken, so it lands where a sequential build would put
        it.  The probe side reads the same table layout either way."""
        from dogqc_amd.hipTranslator import staticLdsBytes, WAVE
        a = self.algExpr
        J = a.opId
        idx = self.ht["index"]
        B, U = pipe.bs, pipe.U
        tt = self.tagType
        tb = 4 if self.tag32 else 8
        words = 1 + len(self.stored)
        fields = ins["fields"]
        prog.hostVars["bpart%d" % idx] = (
            "{ const int32_t req = t->part_build; const uint64_t cap = (uint64_t)L.htcap[%d];\n"
            "  uint32_t np = 0u;\n"
            "  if (req > 0) np = std::min<uint32_t>(1024u, std::max<uint32_t>(2u, pow2floor((uint64_t)req)));\n"
            "  uint64_t span = np ? (cap + np - 1u) / np : 0u;\n"
            "  while (np && np < 1024u && span * %du > %du) { np *= 2u; span = (cap + np - 1u) / np; }\n"
            "  if (span * %du > %du) np = 0u;\n"
            "  L.apart[%d] = np; L.bspan[%d] = np ? (uint32_t)span : 0u;\n"
            "  if (np) {\n"
            "    const uint32_t seg = std::min<uint32_t>(2048u, gridFor(t->grid[%d] > 0 ? t->grid[%d] : %d, %s, %d, 1));\n"
            "    const double mu = (double)std::max<int64_t>(1, %s) / ((double)seg * np);\n"
            "    L.apseg[%d] = seg;\n"
            "    L.apcap[%d] = std::min<int64_t>(((int64_t)1 << 31) / seg, (int64_t)(mu * 1.25 + 4.0 * std::sqrt(mu)) + 16);\n"
            "    L.bspill[%d] = (uint32_t)std::min<uint64_t>(1u << 30, std::max<uint64_t>(4096u, cap / 64u));\n"
            "  } }"
            % (idx, tb, self.PART_BUILD_SPAN_MAX,
               tb, self.PART_BUILD_SPAN_MAX, idx, idx, pipe.pipeId, pipe.pipeId, pipe.defaultGrid,
               pipe.params["n"][1], B * U, est, idx, idx, idx))
        recs = "(uint64_t)L.apart[%d] * L.apseg[%d] * (uint64_t)L.apcap[%d] * %d" % (idx, idx, idx, words)
        fnp = pipe.param("bnp%d" % J, "uint32_t", "L.apart[%d]" % idx)
        fseg = pipe.param("bseg%d" % J, "uint32_t", "L.apseg[%d]" % idx)
        fpcap = pipe.param("bpcap%d" % J, "uint32_t", "L.apcap[%d]" % idx)
        fspan = pipe.param("bspan%d" % J, "uint32_t", "L.bspan[%d]" % idx)
        bcnt = prog.buffer("bpc%d" % J, CType.UINT, "(uint64_t)L.apart[%d] * L.apseg[%d]" % (idx, idx))
        fcnt = pipe.param("bpc%d" % J, "uint32_t*", "B_" + bcnt)
        brec = prog.buffer("bprec%d" % J, CType.UINT64, recs)
        frec = pipe.param("bprec%d" % J, "uint64_t*", "B_" + brec)
        swords = 2 + len(self.stored)
        bsp = prog.buffer("bsp%d" % J, CType.UINT64, "(uint64_t)L.bspill[%d] * %d" % (idx, swords))
        bspc = prog.buffer("bspc%d" % J, CType.UINT, "1", "0x0ull")
        fsp = pipe.param("bsp%d" % J, "uint64_t*", "B_" + bsp)
        fspc = pipe.param("bspc%d" % J, "uint32_t*", "B_" + bspc)
        fspcap = pipe.param("bspcap%d" % J, "uint32_t", "L.bspill[%d]" % idx)
        pipe.annotations.append("PB%d" % J)
        cur, pref, scan = "bcur%d" % J, "bpref%d" % J, "bscan%d" % J
        from dogqc_amd.hipTranslator import AggregationTranslator as AT
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (cur, 1024))
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (pref, AT.PART_SEG_MAX + 1))
        pipe.staticLds.append("__shared__ uint32_t %s[%d];" % (scan, 1024 // WAVE + 1))
        img = "bimg%d" % J
        pro = pipe.prologue
        pro.emit("%s* %s = (%s*)dqc_dyn_lds;" % (tt, img, tt))
        pro.emit("const uint32_t blo%d = std::min<uint32_t>(%s, blockIdx.x * %s);" % (J, ins["fcap"], fspan))
        pro.emit("const uint32_t bhi%d = std::min<uint32_t>(%s, blo%d + %s);" % (J, ins["fcap"], J, fspan))
        # (uniform branch: no barrier when the build is not partitioned)
        pro.open("if (%s > 0u)" % fnp)
        pro.emit("for (uint32_t i = threadIdx.x; i < %s; i += %du) %s[i] = 0u;" % (fnp, B, cur))
        pro.open("if constexpr (PHASE == 1)")
        pro.emit("for (uint32_t i = threadIdx.x; i < bhi%d - blo%d; i += %du) %s[i] = %s[blo%d + i];"
                 % (J, J, B, img, ins["ftags"], J))
        pro.close()
        pro.emit("__syncthreads();")
        pro.close()
        jtag = ins["jtag"]
        need = pipe.reg("bneed%d" % J, "int")
        # PHASE 0: everything before this point, then the insert or the scatter
        pipe.code.lines.insert(0, "    " * (pipe.code.depth - 1) + "if constexpr (PHASE == 0) {")
        pipe.code.depth += 1
        pipe.code.open("if (%s == 0u)" % fnp)
        self._directInsert(pipe, ins, "active[u]")
        pipe.code.close()
        pipe.code.open("else")
        pipe.emit("// append to this workgroup's segment of the row's partition")
        values = [pipe.access(att) for att in self.stored]
        with pipe.rowLoop():
            pipe.emit("%s = 0;" % need)
            pipe.code.open("if (active[u])")
            self._hashCode(pipe, self.buildKeys, jtag, "jh%d" % J, tagIsReg=True)
            if ins["bloom"] is not None:
                fbf, fbfw = ins["bloom"]
                pipe.emit("if (%s) dqc::bloom_set(%s, %s, jh%d);" % (fbfw, fbf, fbfw, J))
            pipe.emit("const uint32_t pb = %s != 0u ? dqc::slot_of(jh%d, %s) / %s : 0xffffffffu;"
                      % (jtag, J, ins["fcap"], fspan))
            pipe.emit("const uint32_t pos = pb != 0xffffffffu ? atomicAdd(&%s[pb], 1u) : 0xffffffffu;" % cur)
            pipe.code.open("if (pos < %s)" % fpcap)
            pipe.emit("uint64_t* rec = %s + (((uint64_t)pb * %s + blockIdx.x) * %s + pos) * %du;"
                      % (frec, fseg, fpcap, words))
            pipe.emit("rec[0] = (uint64_t)%s;" % jtag)
            for k, (att, v) in enumerate(zip(self.stored, values)):
                pipe.emit("rec[%d] = dqc::to_word<%s>(%s);" % (k + 1, pipe.attrs[att.id].ctype, v))
            pipe.code.close()
            pipe.code.open("else")
            pipe.emit("// segment full, or the tag-0 key: insert directly")
            pipe.emit("%s = 1;" % need)
            self._firstCas(pipe, ins)
            pipe.code.close()
            pipe.code.close()
        self._resolveInsert(pipe, ins, "%s" % need, values)
        pipe.code.close()   # end of the scatter
        pipe.code.close()   # end of PHASE 0
        # PHASE 1: the loop index is a record of this workgroup's partition
        pipe.code.open("else")
        rr = pipe.reg("bprec%d" % J, "uint64_t")
        pipe.emit("dqc::part_records<%d>(%s, %s, blockIdx.x, %s, %s, active, %s);"
                  % (U, pref, fseg, fpcap, pipe.tid, rr[:-3]))
        pw = [pipe.reg("bpw%d_%d" % (J, k), "uint64_t") for k in range(len(self.stored))]
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            pipe.emit("const uint64_t* rec = %s + %s * %du;" % (frec, rr, words))
            pipe.emit("%s = (%s)rec[0];" % (jtag, tt))
            for k, r in enumerate(pw):
                pipe.emit("%s = rec[%d];" % (r, k + 1))
            pipe.code.close()
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            if self.tag32:
                pipe.emit("const uint64_t jh = dqc::mix64(dqc::TAG_TOP | (uint64_t)(%s - 1u));" % jtag)
            else:
                pipe.emit("const uint64_t jh = dqc::mix64(%s);" % jtag)
            pipe.emit("const uint32_t hs = dqc::slot_of(jh, %s);" % ins["fcap"])
            pipe.emit("int64_t js = -1;")
            pipe.emit("bool spill = false;")
            pipe.code.open("for (uint32_t q = hs, n = 0;; ++q, ++n)")
            pipe.emit("if (q >= bhi%d) { spill = true; break; }" % J)
            pipe.emit("if (n >= dqc::MAX_PROBE) break;")
            pipe.emit("if (dqc::lds_cas<%s>(&%s[q - blo%d], (%s)0, %s) == (%s)0) { js = (int64_t)q; break; }"
                      % (tt, img, J, tt, jtag, tt))
            pipe.code.close()
            pipe.code.open("if (js >= 0)")
            pipe.emit("%s[js] = %s;" % (ins["ftags"], jtag))
            for f, att, r in zip(fields, self.stored, pw):
                pipe.emit("%s = dqc::from_word<%s>(%s);" % (f.at("js"), pipe.attrs[att.id].ctype, r))
            pipe.code.close()
            pipe.code.open("else if (spill)")
            pipe.emit("const uint32_t sp = atomicAdd(%s, 1u);" % fspc)
            pipe.code.open("if (sp < %s)" % fspcap)
            pipe.emit("uint64_t* srec = %s + (uint64_t)sp * %du;" % (fsp, swords))
            pipe.emit("srec[0] = (uint64_t)%s; srec[1] = hs;" % jtag)
            for k, r in enumerate(pw):
                pipe.emit("srec[%d] = %s;" % (k + 2, r))
            pipe.code.close()
            pipe.emit("else { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); atomicOr(&st->ht_overflow_mask, %du); }"
                      % (1 << idx))
            pipe.code.close()
            pipe.emit("else { dqc::raise_status(&st->flags, DQC_ST_HT_OVERFLOW); atomicOr(&st->ht_overflow_mask, %du); }"
                      % (1 << idx))
            pipe.code.close()
        pipe.code.close()
        # PHASE 0 end: the segment counts
        pipe.epilogue.open("if (PHASE == 0 && %s > 0u)" % fnp)
        pipe.epilogue.emit("__syncthreads();")
        pipe.epilogue.emit("for (uint32_t i = threadIdx.x; i < %s; i += %du) %s[(uint64_t)i * %s + blockIdx.x] = "
                           "std::min<uint32_t>(%s[i], %s);" % (fnp, B, fcnt, fseg, cur, fpcap))
        pipe.epilogue.close()
        pipe.phased = True
        pipe.phasedHt = self.ht
        pipe.phaseBetween = []
        pipe.partAgg = {"np": fnp, "cap": fpcap, "count": fcnt, "nseg": fseg, "pref": pref, "scan": scan,
                        "hostNp": "L.apart[%d]" % idx, "hostNseg": "L.apseg[%d]" % idx}
        pipe.dynLdsExpr = "(uint64_t)L.bspan[%d] * %du" % (idx, tb)
        cols = [(b, ct) for b, ct in self.storedBuffers]

        def spillStep(prog, J=J, idx=idx, tt=tt, swords=swords, cols=cols, bsp=bsp, bspc=bspc):
            lines = ["if (L.apart[%d] > 0u) {" % idx, "dqc::spill_cols sc{};", "sc.n = %du;" % len(cols)]
            for k, (b, ct) in enumerate(cols):
                lines.append("sc.ptr[%d] = (void*)B_%s; sc.size[%d] = (uint32_t)sizeof(%s);" % (k, b, k, ct))
            lines.append("hipLaunchKernelGGL(dqc::dqc_spill_insert<%s>, dim3(64), dim3(256), 0, s, (%s*)B_jht%d_tags, "
                         "(uint32_t)L.htcap[%d], (const uint64_t*)B_%s, %du, (const uint32_t*)B_%s, "
                         "(uint32_t)L.bspill[%d], sc, &st->flags, &st->ht_overflow_mask, %du);"
                         % (tt, tt, J, idx, bsp, swords, bspc, idx, 1 << idx))
            lines.append("}")
            return lines
        pipe.phaseAfter = [{"name": "dqc_spill_insert", "host": spillStep}]

    def _storeEntry(self, pipe, idx, fecap, fields, pos):
        pipe.code.open("if (%s < %s)" % (pos, fecap))
        for f, att in zip(fields, self.stored):
            pipe.emit("%s = %s;" % (f.at(pos), pipe.access(att)))
        pipe.code.close()
        pipe.emit("else { dqc::raise_status(&st->flags, DQC_ST_MM_OVERFLOW); "
                  "atomicOr(&st->mm_overflow_mask, %du); }" % (1 << idx))

    def _multiBuildStaged(self, pipe, J, idx, fdir, fcap, fecap, fields, jb):
        """Count pass that also ranks every row within its bucket, insert
        pass without atomics.

        PHASE 0: rows of one loop trip that fall into the same bucket share
        one LDS counter, so a bucket costs one global atomic per workgroup and
        trip instead of one per row (skewed keys otherwise serialise on one
        address); that atomic returns the group's first rank in the bucket,
        and each row keeps base + its rank in the group (rows without an LDS
        slot take their rank from their own atomic).  The bucket scan keeps
        the counts.  PHASE 1 re-runs the pipeline -- the same rows at the
        same indices -- and stores each entry at bucket offset + rank: no
        atomics and no barriers (the ranked insert replaced one returning
        global atomic per bucket and trip)."""
        W = 1
        while W < pipe.bs * pipe.U:
            W *= 2
        prog = self._prog
        self.ht["ranked"] = True
        rb = prog.buffer("mmr%d" % J, CType.UINT, pipe.params["n"][1])
        frank = pipe.param("mmr%d" % J, "uint32_t*", "B_" + rb)
        pipe.staticLds.append("__shared__ uint32_t mmk%d[%d];" % (J, W))
        pipe.staticLds.append("__shared__ uint32_t mmc%d[%d];" % (J, W))
        pipe.staticLds.append("__shared__ uint32_t mmb%d[%d];" % (J, W))
        pipe.prologue.open("if constexpr (PHASE == 0)")
        pipe.prologue.emit("for (uint32_t i = threadIdx.x; i < %du; i += %du) { mmk%d[i] = dqc::LKEY_EMPTY; "
                           "mmc%d[i] = 0u; }" % (W, pipe.bs, J, J))
        pipe.prologue.emit("__syncthreads();")
        pipe.prologue.close()
        ls = pipe.reg("mmls%d" % J, "int")
        rank = pipe.reg("mmrank%d" % J, "uint32_t")
        tid = pipe.tid + "[u]"
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            self._hashCode(pipe, self.buildKeys, "jtag%d" % J, "jh%d" % J)
            pipe.emit("%s = dqc::slot_of(jh%d, %s);" % (jb, J, fcap))
            pipe.code.close()
        pipe.code.open("if constexpr (PHASE == 0)")
        # bucket, LDS slot and rank within the trip (or the row's own atomic)
        with pipe.rowLoop():
            pipe.emit("%s = -1;" % ls)
            pipe.code.open("if (active[u])")
            pipe.emit("%s = dqc::lds_claim32(mmk%d, %du, %s, 8);" % (ls, J, W - 1, jb))
            pipe.emit("%s = %s >= 0 ? atomicAdd(&mmc%d[%s], 1u) : atomicAdd(&%s[2u * %s + 1u], 1u);"
                      % (rank, ls, J, ls, fdir, jb))
            pipe.code.close()
        # one global atomic per distinct bucket of the trip: the group's base rank
        pipe.emit("__syncthreads();")
        pipe.code.open("for (uint32_t i = threadIdx.x; i < %du; i += %du)" % (W, pipe.bs))
        pipe.emit("const uint32_t k = mmk%d[i];" % J)
        pipe.emit("if (k != dqc::LKEY_EMPTY) mmb%d[i] = atomicAdd(&%s[2u * k + 1u], mmc%d[i]);" % (J, fdir, J))
        pipe.code.close()
        pipe.emit("__syncthreads();")
        with pipe.rowLoop():
            pipe.emit("if (active[u]) %s[%s] = %s >= 0 ? mmb%d[%s] + %s : %s;" % (frank, tid, ls, J, ls, rank, rank))
        pipe.emit("__syncthreads();")
        pipe.emit("for (uint32_t i = threadIdx.x; i < %du; i += %du) { mmk%d[i] = dqc::LKEY_EMPTY; mmc%d[i] = 0u; }"
                  % (W, pipe.bs, J, J))
        pipe.emit("__syncthreads();")
        pipe.code.close()
        # PHASE 1: each entry at its bucket's offset plus its rank
        pipe.code.open("else")
        with pipe.rowLoop():
            pipe.code.open("if (active[u])")
            pipe.emit("const uint32_t jpos = %s[2u * %s] + %s[%s];" % (fdir, jb, frank, tid))
            self._storeEntry(pipe, idx, fecap, fields, "jpos")
            pipe.code.close()
        pipe.code.close()

    # ----------------------------------------------------------- probe
    def _storedFields(self, pipe):
        J = self.algExpr.opId
        if self.ht["kind"] != "unique":
            fe = pipe.param("jht%d_e" % J, "const %s*" % self.entryRecord, "B_jht%d_e" % J)
            return [StoredField(fe, k) for k in range(len(self.stored))]
        if self.packed:
            return [PackedField(pipe.param("jht%d" % J, "const uint64_t*", "B_jht%d_tags" % J), self.packedType)]
        out = []
        for k, (b, ct) in enumerate(self.storedBuffers):
            out.append(StoredField(pipe.param("jht%d_p%d" % (J, k), "const %s*" % ct, "B_" + b)))
        return out

    def _storedType(self, k):
        if self.ht["kind"] != "unique":
            return self.storedTypes[k]
        return self.packedType if self.packed else self.storedBuffers[k][1]

    def _bindBuild(self, pipe, fields, indexReg, nullReg=None):
        """Build-side attributes become lazy loads from the stored columns at
        the row's position ``indexReg[u]``; with ``nullReg`` (outer join) a
        null row reads 0 / the empty string instead."""
        for k, (f, att) in enumerate(zip(fields, self.stored)):
            ct = self._storedType(k)
            load = f.at(indexReg + "[u]")
            if nullReg is not None:
                zero = "dqc::str_t{nullptr, nullptr}" if ct == CType.STR_TYPE else "(%s)0" % ct
                load = "(%s[u] ? %s : %s)" % (nullReg, zero, load)
                deps = [indexReg, nullReg]
            else:
                deps = [indexReg]
            pipe.setLazy(att, load, deps, ct)

    def _matchCond(self, ctxt, pipe, fields, indexReg, uniform=False):
        """C expression: stored keys at indexReg[u] equal the probe keys and
        the extra join conditions hold.  Loads what it reads explicitly.
        With `uniform` (the caller is at a wavefront-uniform position) and
        StringHash.wave, string keys are compared by the whole wavefront
        (emitted here, ahead of the expression)."""
        from dogqc_amd.hipTranslator import StringHash
        a = self.algExpr
        ix = indexReg + "[u]"
        parts = []
        if a.multimatch or not self.exact:
            byId = {att.id: f for f, att in zip(fields, self.stored)}
            for b, p in zip(self.buildKeys, self.probeKeys):
                f = byId[b.id]
                if b.dataType == Type.STRING and uniform and StringHash.wave:
                    name = "weq%d" % pipe.uniqueId()
                    # every lane takes part, but only active lanes' positions
                    # are valid entries: the others pass an empty string
                    pipe.emit("const bool %s = dqc::string_equals_wave(active[u], active[u] ? %s : dqc::str_t{}, %s);"
                              % (name, f.at(ix), pipe.access(p)))
                    parts.append(name)
                elif b.dataType == Type.STRING:
                    parts.append("dqc::string_equals(%s, %s)" % (f.at(ix), pipe.access(p)))
                else:
                    parts.append("(%s == %s)" % (f.at(ix), pipe.access(p)))
        if a.conditions is not None:
            buildIds = set(att.id for att in self.stored)
            condAttrs = []
            from dogqc_amd.hipTranslator import _collectAttrs
            _collectAttrs(a.conditions, condAttrs)
            for att in condAttrs:
                st = pipe.attrs.get(att.id)
                if att.id in buildIds:
                    pipe.emit("%s[u] = %s;" % (st.reg, st.loader() if callable(st.loader) else st.loader))
                    st.loaded = True
            parts.append("(%s)" % ctxt.exprEmitter().emit(a.conditions))
        return " && ".join(parts) if parts else "true"

    def consumeProbe(self, ctxt):
        a = self.algExpr
        pipe = ctxt.pipe
        prog = ctxt.program
        J = a.opId
        idx = self.ht["index"]
        pipe.emit("// hash join probe (opId %d)" % J)
        pipe.need(self.probeKeys)
        condProbe = list(getattr(a, "conditionProbeAttributes", {}).values())
        pipe.need(condProbe)
        if self.ht["kind"] == "unique":
            tt = self.tagType
            wt = self.slotType      # the slot word: uint64_t for packed slots
            ftags = pipe.param("jht%d" % J, "const %s*" % wt, "B_jht%d_tags" % J)
            fcap = pipe.param("jht%d_cap" % J, "uint32_t", "L.htcap[%d]" % idx)
            fields = self._storedFields(pipe)
            slotName = "jslot%d" % J
            slot = pipe.reg(slotName, "int64_t", "-1")
            pipe.indexRegs[slotName] = "int64_t"
            self._bindBuild(pipe, fields, slotName)
            needCheck = (not self.exact) or a.conditions is not None
            jtag = pipe.reg("jtag%d" % J, tt)
            jc = pipe.reg("jc%d" % J, "uint32_t")
            jw = pipe.reg("jw%d" % J, "dqc::probe_win<%s>" % wt)
            lbf = self._probeFilter(ctxt, pipe, idx)
            jf = pipe.reg("jf%d" % J, "int") if lbf is not None else None
            # in LDS when it fits (p.jbfpJ words), else read where it lies in
            # device memory, where it stays in L2 (p.jbfgJ words), else none
            passes = ("(p.jbfp%d ? dqc::bloom_test(%s, p.jbfp%d, jh%d) : "
                      "dqc::bloom_test(p.jbf%d, p.jbfg%d, jh%d))" % (J, lbf, J, J, J, J, J)) \
                if lbf is not None else None
            if not needCheck:
                # the probe window of every row of the thread first
                # (independent loads), then the comparisons
                with pipe.rowLoop():
                    pipe.emit("%s = -1;" % slot)
                    pre = self._preHash(pipe)
                    pipe.code.open("if (active[u])")
                    self._hashCode(pipe, self.probeKeys, jtag, "jh%d" % J, tagIsReg=True, pre=pre)
                    pipe.emit("%s = dqc::probe_start(jh%d, %s);" % (jc, J, fcap))
                    if lbf is not None:
                        # rows the filter rules out load no probe window
                        pipe.emit("%s = (p.jbfp%d == 0u && p.jbfg%d == 0u) || %s;" % (jf, J, J, passes))
                        pipe.emit("if (%s) %s = dqc::probe_window(%s, %s);" % (jf, jw, ftags, jc))
                    else:
                        pipe.emit("%s = dqc::probe_window(%s, %s);" % (jw, ftags, jc))
                    pipe.code.close()
            with pipe.rowLoop():
                if needCheck:
                    pipe.emit("%s = -1;" % slot)
                pre = self._preHash(pipe) if needCheck else None
                pipe.code.open("if (active[u])")
                if needCheck:
                    self._hashCode(pipe, self.probeKeys, jtag, "jh%d" % J, tagIsReg=True, pre=pre)
                    if lbf is not None:
                        pipe.emit("%s = (p.jbfp%d == 0u && p.jbfg%d == 0u) || %s;" % (jf, J, J, passes))
                        pipe.code.open("if (%s)" % jf)
                    pipe.emit("uint32_t jcur = dqc::probe_start(jh%d, %s), jv = 0;" % (J, fcap))
                    if self.tag32:
                        # tag 0: only the extra slot can hold the key
                        pipe.code.open("if (%s == 0u)" % jtag)
                        pipe.emit("%s = %s[%s] != 0u ? (int64_t)%s : -1;" % (slot, ftags, fcap, fcap))
                        pipe.emit("if (%s >= 0 && !(%s)) %s = -1;"
                                  % (slot, self._matchCond(ctxt, pipe, fields, slotName), slot))
                        pipe.code.close()
                        pipe.code.open("else for (;;)")
                    else:
                        pipe.code.open("for (;;)")
                    pipe.emit("%s = %s(%s, %s, %s, jcur, jv);" % (slot, self._probeFn(), ftags, fcap, jtag))
                    pipe.emit("if (%s < 0) break;" % slot)
                    pipe.emit("if (%s) break;" % self._matchCond(ctxt, pipe, fields, slotName))
                    pipe.code.close()
                    if lbf is not None:
                        pipe.code.close()
                else:
                    if lbf is not None:
                        pipe.code.open("if (%s)" % jf)
                    pipe.emit("uint32_t jcur = 0, jv = 0;")
                    pipe.emit("%s = %s(%s, %s, %s, %s, jcur, jv);"
                              % (slot, "dqc::probe_in_window_packed" if self.packed else
                                 "dqc::probe_in_window<%s>" % tt, jw, jc, fcap, jtag))
                    pipe.emit("if (%s == -2) %s = %s(%s, %s, %s, jcur, jv);"
                              % (slot, slot, self._probeFn(), ftags, fcap, jtag))
                    if self.tag32:
                        pipe.emit("if (%s == 0u) %s = %s[%s] != 0u ? (int64_t)%s : -1;"
                                  % (jtag, slot, ftags, fcap, fcap))
                    if lbf is not None:
                        pipe.code.close()
                if a.joinType == Join.ANTI:
                    pipe.emit("active[u] = %s < 0;" % slot)
                else:
                    pipe.emit("active[u] = %s >= 0;" % slot)
                pipe.code.close()
                if a.joinType == Join.INNER:
                    for b in self.keyReconstruct:
                        p = self.probeKeys[self.buildKeys.index(b)]
                        st = pipe.setComputed(b)
                        pipe.emit("if (active[u]) %s[u] = %s;" % (st.reg, pipe.access(p)))
            if a.joinType == Join.INNER and self.exact and not self.packed:
                # a group-by on this join's key can accumulate at the join slot
                pipe.probedJoins.append({"join": self, "slot": slotName, "htIndex": idx, "opId": J,
                                         "loopOps": pipe.loopOps,
                                         "buildKeyIds": [b.id for b in self.buildKeys], "tag32": self.tag32,
                                         "probeKeyIds": [p.id for p in self.probeKeys]})
            self.parent.consume(ctxt)
            return

        fdir = pipe.param("jht%d_dir" % J, "const uint32_t*", "B_jht%d_dir" % J)
        fcap = pipe.param("jht%d_cap" % J, "uint32_t", "L.htcap[%d]" % idx)
        # entries past the capacity were never written (the insert pass
        # flagged an overflow and the run is repeated): clamp every range
        self.fecap = pipe.param("jht%d_ecap" % J, "uint32_t", "L.mment[%d]" % idx)
        fields = self._storedFields(pipe)
        posName = "mpos%d" % J
        pos = pipe.reg(posName, "uint32_t", "0")
        pipe.indexRegs[posName] = "uint32_t"
        if a.joinType in (Join.SEMI, Join.ANTI):
            with pipe.rowLoop():
                pre = self._preHash(pipe)
                pipe.code.open("if (active[u])")
                self._hashCode(pipe, self.probeKeys, "jtag%d" % J, "jh%d" % J, pre=pre)
                pipe.emit("const uint2 jd = ((const uint2*)%s)[dqc::slot_of(jh%d, %s)];" % (fdir, J, fcap))
                pipe.emit("int fm = 0;")
                pipe.emit("const uint32_t jend = (uint32_t)std::min<uint64_t>((uint64_t)jd.x + jd.y, %s);"
                          % self.fecap)
                pipe.code.open("for (%s = jd.x; %s < jend; ++%s)" % (pos, pos, pos))
                self._bindBuild(pipe, fields, posName)
                pipe.emit("if (%s) { fm = 1; break; }" % self._matchCond(ctxt, pipe, fields, posName))
                pipe.code.close()
                pipe.emit("active[u] = %s;" % ("fm" if a.joinType == Join.SEMI else "!fm"))
                pipe.code.close()
            self.parent.consume(ctxt)
            return
        if pipe.U != 1:
            raise RuntimeError("multimatch probes need a one-row-per-thread pipeline")
        probeSide = self.rightChild.algExpr.outRelation
        probeLive = [att for id, att in a.outRelation.items() if id in probeSide]
        split = self._splitEnabled(pipe)
        pipe.loopOps += 1
        if a.joinType == Join.OUTER:
            # the probe row's own lane tracks whether it found a match, so
            # outer joins walk lists per lane
            self._probeLoop(ctxt, pipe, fdir, fcap, fields, posName, probeLive, outer=True)
        elif EquiJoinTranslator.usePushDownJoin and ctxt.innerLoopCount == 0:
            self._probePushDown(ctxt, pipe, fdir, fcap, fields, posName, probeLive, split)
        else:
            self._probeLoop(ctxt, pipe, fdir, fcap, fields, posName, probeLive)

    def _probeFilter(self, ctxt, pipe, idx):
        """The build's filter in this pipeline's LDS (its first filtered join
        only), or None.  The pipeline copies it at its start; its word count
        p.jbfpJ is 0 when it does not fit beside the pipeline's other LDS."""
        if not self.filtered or pipe.ldsFilter is not None:
            return None
        from dogqc_amd.hipTranslator import staticLdsBytes
        J = self.algExpr.opId
        prog = ctxt.program
        pid = pipe.pipeId
        # words in this pipeline: the build's, if they fit with the static LDS
        # and 16 KB to spare for an LDS group table
        prog.hostVars["bfp%d" % pid] = lambda: (
            "L.bfp[%d] = (uint64_t)L.bfw[%d] * 4u + %du + 16384u <= maxLds() ? L.bfw[%d] : 0u;"
            % (pid, idx, staticLdsBytes(pipe) + 1024, idx))
        fbf = pipe.param("jbf%d" % J, "const uint32_t*", "B_jbf%d" % J)
        pipe.param("jbfp%d" % J, "uint32_t", "L.bfp[%d]" % pid)
        pipe.param("jbfg%d" % J, "uint32_t", "L.bfp[%d] ? 0u : L.bfw[%d]" % (pid, idx))
        lbf = "lbf%d" % J
        pipe.prologue.emit("uint32_t* %s = (uint32_t*)dqc_dyn_lds;" % lbf)
        pipe.prologue.emit("for (uint32_t i = threadIdx.x; i < p.jbfp%d; i += %du) %s[i] = %s[i];"
                           % (J, pipe.bs, lbf, fbf))
        pipe.prologue.emit("__syncthreads();")
        pipe.ldsFilter = {"dev": "(uint64_t)p.jbfp%d * 4u" % J, "host": "(uint64_t)L.bfp[%d] * 4u" % pid,
                          "param": "jbfp%d" % J}
        if pipe.dynLdsExpr is None:
            pipe.dynLdsExpr = pipe.ldsFilter["host"]
        pipe.annotations.append("BF%d" % J)
        return lbf

    def _probeLoop(self, ctxt, pipe, fdir, fcap, fields, posName, probeLive, outer=False):
        """Each lane walks its own match list; the wavefront loops while any
        lane has matches left (hashJoins.py:465-534).  An outer join gives a
        probe row without a match one more trip as a null row: its build
        attributes read 0 / empty and carry a null flag that aggregations and
        materialisation honour (hashJoins.py:473-530; cudaTranslator.py:475)."""
        J = self.algExpr.opId
        moff = pipe.reg("moff%d" % J, "uint32_t", "0")
        mend = pipe.reg("mend%d" % J, "uint32_t", "0")
        pos = posName + "[u]"
        pipe.emit("%s = 0; %s = 0;" % (moff, mend))
        pre = self._preHash(pipe)
        pipe.code.open("if (active[u])")
        self._hashCode(pipe, self.probeKeys, "jtag%d" % J, "jh%d" % J, pre=pre)
        pipe.emit("const uint2 jd = ((const uint2*)%s)[dqc::slot_of(jh%d, %s)];" % (fdir, J, fcap))
        pipe.emit("%s = jd.x; %s = (uint32_t)std::min<uint64_t>((uint64_t)jd.x + jd.y, %s);"
                  % (moff, mend, self.fecap))
        pipe.code.close()
        if not outer:
            # an empty bucket ends an inner probe row here, so the probe-side
            # columns read below are not fetched for it
            pipe.emit("if (%s >= %s) active[u] = 0;" % (moff, mend))
        pipe.need(probeLive)
        pipe.emit("int mact%d = active[u] && %s < %s;" % (J, moff, mend))
        nullReg = None
        if outer:
            nullReg = "onl%d" % J
            pipe.reg(nullReg, "int", "0")
            pipe.indexRegs[nullReg] = "int"
            opend = "opend%d" % J
            pipe.emit("int %s = active[u];   // probe row still owes its null row" % opend)
            pipe.code.open("while (dqc::ballot(mact%d || %s))" % (J, opend))
            pipe.emit("%s[u] = !mact%d && %s;" % (nullReg, J, opend))
            pipe.emit("%s = mact%d ? %s : 0u;" % (pos, J, moff))
            for att in self.stored:
                pipe.nullFlags[att.id] = nullReg + "[u]"
        else:
            pipe.code.open("while (dqc::ballot(mact%d))" % J)
            pipe.emit("%s = %s;" % (pos, moff))
        self._bindBuild(pipe, fields, posName, nullReg)
        pipe.emit("active[u] = mact%d;" % J)
        pipe.emit("if (active[u]) active[u] = (%s) ? 1 : 0;" % self._matchCond(ctxt, pipe, fields, posName,
                                                                                 uniform=True))
        if outer:
            pipe.emit("if (active[u]) %s = 0;" % opend)
            pipe.emit("if (%s[u]) { active[u] = 1; %s = 0; }" % (nullReg, opend))
        ctxt.innerLoopCount += 1
        pipe.innerDepth += 1
        self.parent.consume(ctxt)
        pipe.innerDepth -= 1
        ctxt.innerLoopCount -= 1
        pipe.emit("++%s;" % moff)
        pipe.emit("mact%d = mact%d && %s < %s;" % (J, J, moff, mend))
        pipe.code.close()
        pipe.emit("active[u] = 0;")

    def _splitEnabled(self, pipe):
        """Long-list splitting needs the pipeline's first expansion (no lane
        reassignment before it), no refill marks, and a pipeline that is not
        itself launched in two phases (a multimap build as its sink)."""
        from dogqc_amd.hipTranslator import AggregationTranslator, MaterializeTranslator
        if pipe.innerDepth or pipe.loopOps or pipe.phased or any(l is None for l in pipe.code.lines):
            return False
        t = self
        while t.parent is not None:
            p = t.parent
            if isinstance(p, EquiJoinTranslator) and p.leftChild is t and p.algExpr.multimatch:
                return False
            if isinstance(p, (AggregationTranslator, MaterializeTranslator)):
                break
            t = p
        return True

    def _splitQueue(self, ctxt, pipe, moff, mlen, live):
        """PHASE 0: queue the probe rows whose match list is long; PHASE 1:
        every lane takes one chunk of a queued list (see module docstring)."""
        a = self.algExpr
        J = a.opId
        prog = ctxt.program
        idx = self.ht["index"]
        CH = self.SPLIT_CHUNK
        est = prog.estimate(self.rightChild.algExpr)
        prog.splitVars["scap%d" % idx] = (
            "L.scap[%d] = t->split_capacity[%d] > 0 ? t->split_capacity[%d] : std::max<int64_t>(1024, %s / 256);"
            % (idx, idx, idx, est))
        cap = "L.scap[%d]" % idx
        fcap = pipe.param("scap%d" % J, "uint64_t", cap)
        fthr = pipe.param("sthr%d" % J, "uint32_t", "t->split_threshold[%d] > 0 ? t->split_threshold[%d] : %d"
                          % (pipe.pipeId, pipe.pipeId, self.DEFAULT_SPLIT_THRESHOLD))
        blen = prog.buffer("sq%d_len" % J, CType.UINT, cap)
        boff = prog.buffer("sq%d_off" % J, CType.UINT, cap)
        flen = pipe.param("sq%d_len" % J, "uint32_t*", "B_" + blen)
        foff = pipe.param("sq%d_off" % J, "uint32_t*", "B_" + boff)
        regs = []
        for k, (r, ct) in enumerate(live.items()):
            b = prog.buffer("sq%d_r%d" % (J, k), ct, cap)
            regs.append((r, pipe.param("sq%d_r%d" % (J, k), "%s*" % ct, "B_" + b)))
        bcs = prog.buffer("sq%d_cstart" % J, CType.UINT64, cap + " + 1")
        btot = prog.buffer("sq%d_total" % J, CType.UINT64, "1")
        fcs = pipe.param("sq%d_cstart" % J, "const uint64_t*", "B_" + bcs)
        ftot = pipe.param("sq%d_total" % J, "const uint64_t*", "B_" + btot)
        # PHASE 0: queue long lists (wave-aggregated append), walk the rest
        pipe.code.open("if (active[u] && %s > %s)" % (mlen, fthr))
        pipe.emit("const uint64_t qm = dqc::ballot(true);")
        pipe.emit("const int ql = dqc::first_lane(qm);")
        pipe.emit("unsigned long long qb = 0;")
        pipe.emit("if ((int)lane == ql) qb = atomicAdd((unsigned long long*)&st->split_count[%d], "
                  "(unsigned long long)dqc::popc64(qm));" % idx)
        pipe.emit("const uint64_t qp = dqc::shfl(qb, ql) + dqc::lanes_below(qm);")
        pipe.code.open("if (qp < %s)" % fcap)
        pipe.emit("%s[qp] = (uint32_t)%s;" % (flen, mlen))
        pipe.emit("%s[qp] = %s;" % (foff, moff))
        for r, f in regs:
            pipe.emit("%s[qp] = %s[u];" % (f, r))
        pipe.code.close()
        pipe.emit("%s = 0;" % mlen)
        pipe.code.close()
        pipe.code.close()   # end of PHASE 0
        # PHASE 1: the loop index is a chunk of a queued list
        pipe.code.open("else")
        pipe.emit("const uint64_t hc = (uint64_t)%s[u];" % pipe.tid)
        pipe.emit("%s = 0; %s = 0;" % (moff, mlen))
        pipe.code.open("if (active[u])")
        pipe.emit("const uint64_t hn = std::min<uint64_t>(st->split_count[%d], %s);" % (idx, fcap))
        pipe.emit("const uint32_t hi = dqc::split_item(%s, hn, hc);" % fcs)
        pipe.emit("const uint64_t hk = (hc - %s[hi]) * %du;" % (fcs, CH))
        for r, f in regs:
            pipe.emit("%s[u] = %s[hi];" % (r, f))
        pipe.emit("%s = %s[hi] + (uint32_t)hk;" % (moff, foff))
        pipe.emit("%s = std::min<uint64_t>(%du, (uint64_t)%s[hi] - hk);" % (mlen, CH, flen))
        pipe.code.close()
        pipe.code.close()
        pipe.phased = True
        pipe.phaseRows[1] = ("(int64_t)%s * %d" % (cap, max(64, 4096 // CH)), "*%s" % ftot)

        def scan(prog, idx=idx, blen=blen, bcs=bcs, btot=btot, cap=cap, CH=CH):
            return ["hipLaunchKernelGGL(dqc::dqc_split_scan, dim3(1), dim3(dqc::SCAN_BS), 0, s, "
                    "(const uint32_t*)B_%s, (const uint64_t*)&st->split_count[%d], (uint64_t)%s, %du, "
                    "B_%s, B_%s);" % (blen, idx, cap, CH, bcs, btot)]
        pipe.phaseBetween = [{"name": "dqc_split_scan", "host": scan}]
        pipe.annotations.append("SPLIT%d" % J)

    def _probePushDown(self, ctxt, pipe, fdir, fcap, fields, posName, probeLive, split):
        """Push-down parallelism for wave64 (see module docstring); `split`
        queues long lists for the second launch."""
        J = self.algExpr.opId
        nw = pipe.bs // 64
        pipe.staticLds.append("__shared__ uint64_t pdincl%d[%d][64];" % (J, nw))
        pipe.annotations.append("PD%d" % J)
        moff = pipe.reg("moff%d" % J, "uint32_t", "0")
        mlen = pipe.reg("mlen%d" % J, "uint64_t", "0")
        pos = posName + "[u]"
        if split:
            # everything before this point is the first pass (PHASE 0)
            pipe.code.lines.insert(0, "    " * (pipe.code.depth - 1) + "if constexpr (PHASE == 0) {")
            pipe.code.depth += 1
        pipe.emit("%s = 0; %s = 0;" % (moff, mlen))
        pre = self._preHash(pipe)
        pipe.code.open("if (active[u])")
        self._hashCode(pipe, self.probeKeys, "jtag%d" % J, "jh%d" % J, pre=pre)
        pipe.emit("const uint2 jd = ((const uint2*)%s)[dqc::slot_of(jh%d, %s)];" % (fdir, J, fcap))
        pipe.emit("%s = jd.x; %s = jd.x < %s ? std::min<uint64_t>(jd.y, %s - jd.x) : 0;"
                  % (moff, mlen, self.fecap, self.fecap))
        pipe.code.close()
        # an empty bucket ends the (inner) probe row here, so the probe-side
        # columns read below are not fetched for it
        pipe.emit("if (%s == 0) active[u] = 0;" % mlen)
        pipe.need(probeLive)
        # everything the matching and the operators above read from the probe
        # tuple moves with it to the lanes that process its matches
        condProbe = list(getattr(self.algExpr, "conditionProbeAttributes", {}).values())
        live = pipe.liveState(probeLive + self.probeKeys + condProbe)
        if split:
            self._splitQueue(ctxt, pipe, moff, mlen, live)
        pipe.emit("const uint64_t pincl%d = dqc::wave_incl_scan(%s);" % (J, mlen))
        pipe.emit("const uint64_t ptotal%d = dqc::uniform64(dqc::shfl(pincl%d, 63));" % (J, J))
        pipe.emit("pdincl%d[wave][lane] = pincl%d;" % (J, J))
        saved = []
        for r, ct in live.items():
            sv = pipe.reg("%s_pd%d" % (r, J), ct)
            pipe.emit("%s = %s[u];" % (sv, r))
            saved.append((r, sv))
        pipe.emit("const uint32_t moffs%d = %s;" % (J, moff))
        pipe.code.open("for (uint64_t pdone = 0; pdone < ptotal%d; pdone += 64)" % J)
        pipe.emit("const uint64_t pt = pdone + lane;")
        pipe.emit("active[u] = pt < ptotal%d;" % J)
        pipe.emit("int po = 0;")
        pipe.code.open("if (active[u])")
        pipe.emit("int lo = 0, hi = 63;")
        pipe.emit("while (lo < hi) { const int mid = (lo + hi) >> 1; "
                  "if (pdincl%d[wave][mid] > pt) hi = mid; else lo = mid + 1; }" % J)
        pipe.emit("po = lo;")
        pipe.code.close()
        pipe.emit("const uint64_t pbase = po ? pdincl%d[wave][po - 1] : 0;" % J)
        pipe.emit("%s = dqc::shfl(moffs%d, po) + (uint32_t)(pt - pbase);" % (pos, J))
        for r, sv in saved:
            pipe.emit("%s[u] = dqc::shfl(%s, po);" % (r, sv))
        self._bindBuild(pipe, fields, posName)
        pipe.emit("if (active[u]) active[u] = (%s) ? 1 : 0;" % self._matchCond(ctxt, pipe, fields, posName,
                                                                                 uniform=True))
        ctxt.innerLoopCount += 1
        pipe.innerDepth += 1
        self.parent.consume(ctxt)
        pipe.innerDepth -= 1
        ctxt.innerLoopCount -= 1
        pipe.code.close()
        pipe.emit("active[u] = 0;")
