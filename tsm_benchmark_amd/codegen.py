#!/usr/bin/env python3
"""Generator of the gfx950 HIP kernels for C = A^T B.

This is the counterpart of the reference's ``sycl_codegen_new.py``: it
enumerates kernel variants and emits concrete kernels for them.  Where the
reference prints SYCL kernels (one per M, N, TM, TN and variant,
``sycl_codegen_new.py:142-363``) this prints HIP translation units:

* ``csrc/generated/tsm_kernels_<dtype>_<KC>x<MC>.hip`` - one ``__global__`` kernel per
  configuration (named ``tsm_atb_<dtype>_<KC>x<MC>_t<TM>x<TN>_r<ROWS>_p<PF>``,
  the name rocprof shows) plus its host launcher;
* ``csrc/generated/tsm_table.cpp`` - the kernel table the C ABI dispatches
  through and the tuned default per (dtype, width class), baked in from the
  autotuner's JSON tile table (``tuning/gfx950.json``);
* ``include/tsm/tsmttsm_kernels.hh`` (``--name-layer``) - the reference's
  kernel names ``tsmttsm_{M}_{N}_{TM}_{TN}[_lf|_sg|_sg_lf]`` as thin inline
  wrappers over the C ABI, for the M/TM set given on the command line.

Configuration axes (all compile-time in the emitted kernels):
  dtype    f64 | f32 | bf16      storage type of A and B
  KC, MC   width class           k <= KC, m <= MC, both powers of two <= 64
  TM, TN   register tile         C entries owned by one thread (``TM``/``TN``)
  ROWS     rows per LDS tile     the n-tile streamed per pipeline stage
  PF       tiles in flight       1 or 2 (the reference's "leap-frog" depth)
  NT       load cache policy     0 normal, 1 non-temporal (streaming) loads

CLI flags keep the reference's vocabulary where it maps
(``sycl_codegen_new.py:588-604``): ``--M``, ``--tile-sizes``, ``--stride``,
``--local-size``, ``--sg-size``, ``--unroll``.  On gfx950 the work-group size
is fixed at 256 and the sub-group (wavefront) size at 64; other values are
rejected, as the reference rejects inconsistent flags (``:606-620``).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GEN_DIR = os.path.join(HERE, "csrc", "generated")
TUNING_JSON = os.path.join(HERE, "tuning", "gfx950.json")

DTYPES = ("f64", "f32", "bf16")
DTYPE_ID = {"f64": 0, "f32": 1, "bf16": 2}
ELEM_BYTES = {"f64": 8, "f32": 4, "bf16": 2}
CXX_TYPE = {"f64": "double", "f32": "float", "bf16": "tsm::bf16"}
ARCH = "gfx950"  # the one target of the generated kernels
CLASSES = (1, 2, 4, 8, 16, 32, 64)
THREADS = 256          # work-group size (4 wave64)
WAVE = 64              # gfx950 wavefront size
LDS_BUDGET = 65536     # LDS per work-group of the defaults (two fit in 160 KiB)
LDS_MAX = 131072       # LDS per work-group of any variant (one per compute unit)
DIRECT_DEPTHS = (4, 8, 16)  # steps in flight per wave of the direct engine (FP64, FP32)
DIRECT_DEPTHS_BF16 = (2, 4)  # BF16: even, dividing the 4-step (128-row) fold
MAX_ROWS_PER_THREAD = 16  # the fully unrolled row loop of one tile (compile time)


def lcm(a: int, b: int) -> int:
    return a * b // math.gcd(a, b)


def width_class(x: int) -> int:
    """Smallest class (power of two) >= x, for 1 <= x <= 64."""
    if not 1 <= x <= 64:
        raise ValueError(f"width {x} outside [1, 64]")
    c = 1
    while c < x:
        c *= 2
    return c


@dataclasses.dataclass(frozen=True, order=True)
class KernelConfig:
    dtype: str
    kc: int
    mc: int
    tm: int
    tn: int
    rows: int
    pf: int
    nt: int = 0  # 1: non-temporal (streaming) global loads
    sym: int = 0  # 1: C = A^T A (B is A; only A's tile is staged)
    # 1: matrix cores fed straight from global memory, no LDS staging
    # (rows = 4, the K-step; pf = steps in flight per wave)
    direct: int = 0

    @property
    def mfma(self) -> bool:
        """tm = tn = 0 selects the matrix-core consumer."""
        return self.tm == 0 and self.tn == 0

    @property
    def engine(self) -> int:
        """1 VALU register tiles, 2 matrix cores through LDS, 3 direct matrix cores."""
        return 3 if self.direct else (2 if self.mfma else 1)

    @property
    def name(self) -> str:
        op = "ata" if self.sym else "atb"
        if self.direct:
            return (f"tsm_{op}_{self.dtype}_{self.kc}x{self.mc}_direct_d{self.pf}"
                    + ("_nt" if self.nt else ""))
        tile = "mfma" if self.mfma else f"t{self.tm}x{self.tn}"
        return (f"tsm_{op}_{self.dtype}_{self.kc}x{self.mc}_{tile}"
                f"_r{self.rows}_p{self.pf}" + ("_nt" if self.nt else ""))

    @property
    def kstep(self) -> int:
        """Rows per matrix instruction (MFMA consumer)."""
        return 32 if self.dtype == "bf16" else 4

    @property
    def ntiles(self) -> int:
        return 1 if self.mfma else (self.kc // self.tm) * (self.mc // self.tn)

    @property
    def row_lanes(self) -> int:
        return THREADS // self.ntiles

    @property
    def kp(self) -> int:
        """Width of A the consumer works on: the staged matrix-core consumer
        takes 16-wide blocks, so an 8-wide side is staged into LDS rows of 16
        (the upper half zero)."""
        return 16 if self.mfma and not self.direct and self.kc < 16 else self.kc

    @property
    def mp(self) -> int:
        """Width of B the consumer works on (as kp)."""
        return 16 if self.mfma and not self.direct and self.mc < 16 else self.mc

    @property
    def stage_bytes(self) -> int:
        return self.rows * (self.kp + (0 if self.sym else self.mp)) * ELEM_BYTES[self.dtype]

    @property
    def direct_depths(self) -> tuple:
        return DIRECT_DEPTHS_BF16 if self.dtype == "bf16" else DIRECT_DEPTHS

    @property
    def lds_bytes(self) -> int:
        if self.direct and self.dtype == "bf16":  # four waves' double-buffered 32-row steps
            staging = 4 * 2 * 32 * (self.kc + (0 if self.sym else self.mc)) * 2
            return max(staging, self.kc * self.mc * 8, THREADS * 16)
        if self.direct:  # two KC x MC FP64 halves of the cross-wave sum
            return max(2 * self.kc * self.mc * 8, THREADS * 16)
        return max(2 * self.stage_bytes, self.kp * self.mp * 8, THREADS * 8)

    def problems(self) -> list[str]:
        """Reasons this configuration cannot be emitted (empty when valid)."""
        p = []
        es = ELEM_BYTES[self.dtype]
        if self.dtype not in DTYPES:
            p.append("dtype")
        if self.kc not in CLASSES or self.mc not in CLASSES:
            p.append("width class must be a power of two <= 64")
            return p
        if self.sym not in (0, 1) or (self.sym and self.kc != self.mc):
            p.append("A^T A kernels are square")
        if self.direct:
            if self.direct != 1:
                p.append("direct must be 0 or 1")
            if self.dtype not in DTYPES:
                p.append("dtype")
            if self.kc % 16 or self.mc % 16:
                p.append("MFMA needs width classes that are multiples of 16")
            if self.tm or self.tn:
                p.append("the direct engine has no register tile")
            if self.rows != self.kstep:
                p.append(f"direct-engine steps are {self.kstep} rows")
            if self.pf not in self.direct_depths or self.nt not in (0, 1):
                p.append(f"steps in flight must be one of {self.direct_depths}, nt 0 or 1")
            if self.lds_bytes > LDS_MAX:
                p.append("LDS budget")
            return p
        if self.mfma:
            # a narrower side whose rows are whole 16-byte chunks (or 8 bytes,
            # two rows to a chunk) is padded to 16 in LDS: beside a side of 16
            # or more, or for A^T A at 8 (one diagonal block from one padded
            # tile)
            if (any(w % 16 and (w * es) % 16 and w * es != 8 for w in (self.kc, self.mc))
                    or (max(self.kc, self.mc) < 16 and not (self.sym and self.kc == 8))):
                p.append("MFMA needs width classes that are multiples of 16 (or a narrower "
                         "side of whole 16-byte or 8-byte rows beside one)")
            if self.rows % (4 * self.kstep):
                p.append(f"MFMA rows must be a multiple of {4 * self.kstep}")
            if self.rows // (4 * self.kstep) > MAX_ROWS_PER_THREAD:
                p.append("too many MFMA steps per tile")
            if self.pf not in (1, 2) or self.nt not in (0, 1):
                p.append("pf must be 1 or 2, nt 0 or 1")
            if (self.rows * self.kc * es) % 16 or (self.rows * self.mc * es) % 16:
                p.append("tiles must be whole 16-byte chunks")
            if self.lds_bytes > LDS_MAX:
                p.append("LDS budget")
            return p
        if self.tm < 1 or self.tn < 1 or self.kc % self.tm or self.mc % self.tn:
            p.append("tile must divide the width class")
            return p
        if self.ntiles > WAVE:
            p.append("more than 64 register tiles")
        if self.pf not in (1, 2, 3, 4):
            p.append("pf must be 1 to 4")
        if self.nt not in (0, 1):
            p.append("nt must be 0 or 1")
        if self.rows < 1 or self.rows % self.row_lanes:
            p.append(f"rows must be a multiple of {self.row_lanes} row lanes")
        if (self.rows * self.kc * es) % 16 or (self.rows * self.mc * es) % 16:
            p.append("tiles must be whole 16-byte chunks")
        if self.lds_bytes > LDS_MAX:
            p.append("LDS budget")
        if self.rows // max(self.row_lanes, 1) > MAX_ROWS_PER_THREAD:
            p.append(f"more than {MAX_ROWS_PER_THREAD} rows per thread per tile")
        return p


def _pow2_floor(x: int) -> int:
    return 1 << (max(1, x).bit_length() - 1)


def default_config(dtype: str, kc: int, mc: int, *, pf: int = 2, nt: int = 1) -> KernelConfig:
    """Heuristic configuration used when the tile table has no entry.

    FP64 keeps up to an 8 x 8 register tile (64 FP64 accumulators); FP32/BF16
    keep up to 8 x 4 because each entry also carries a (hi, lo) pair.  The
    row tile aims at 32 KiB per pipeline stage."""
    es = ELEM_BYTES[dtype]
    if kc % 16 == 0 and mc % 16 == 0 and (dtype == "bf16" or kc * mc >= 1024):
        # matrix cores: the register tiles would be VALU-bound at these widths
        rows = _pow2_floor(32768 // ((kc + mc) * es))
        c = KernelConfig(dtype, kc, mc, 0, 0, max(rows, 4 * (32 if dtype == "bf16" else 4)), pf, nt)
        while c.lds_bytes > LDS_BUDGET and c.rows > 4 * c.kstep:
            c = dataclasses.replace(c, rows=c.rows // 2)
        while c.rows // (4 * c.kstep) > MAX_ROWS_PER_THREAD:
            c = dataclasses.replace(c, rows=c.rows // 2)
        if not c.problems():
            return c
    tm = min(kc, 8)
    tn = min(mc, 8 if dtype == "f64" else 4)
    if (kc // tm) * (mc // tn) > WAVE:
        tn = min(mc, 8)
    ntiles = (kc // tm) * (mc // tn)
    rg = THREADS // ntiles
    rows = _pow2_floor(32768 // ((kc + mc) * es))
    rows = max(rows, rg)
    while (rows * kc * es) % 16 or (rows * mc * es) % 16:
        rows *= 2
    while rows // rg > MAX_ROWS_PER_THREAD and rows > rg:
        rows //= 2
    cfg = KernelConfig(dtype, kc, mc, tm, tn, rows, pf, nt)
    while cfg.lds_bytes > LDS_BUDGET and cfg.rows > rg:
        cfg = dataclasses.replace(cfg, rows=cfg.rows // 2)
    bad = cfg.problems()
    if bad:
        raise ValueError(f"no default for {dtype} {kc}x{mc}: {bad}")
    return cfg


# Shapes of the sweep / precision / batched workloads that get extra variants
# for the autotuner to choose from.
TUNE_SHAPES = {
    "f64": [(1, 1), (2, 2), (4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 32), (32, 4)],
    "f32": [(8, 8), (16, 16), (32, 32), (64, 64)],
    "bf16": [(8, 8), (16, 16), (32, 32), (64, 64)],
}


def tuning_variants(dtype: str, kc: int, mc: int) -> list[KernelConfig]:
    """Variants the autotuner chooses from around the heuristic default: the
    register tile, rows per tile and pipeline depth, plus non-temporal loads
    for the default tile."""
    base = default_config(dtype, kc, mc)
    out = set()
    if kc % 16 == 0 and mc % 16 == 0:
        mrows = _pow2_floor(32768 // ((kc + mc) * ELEM_BYTES[dtype]))
        for rows in (mrows // 2, mrows, mrows * 2):
            for pf in (1, 2):
                for nt in (0, 1):
                    c = KernelConfig(dtype, kc, mc, 0, 0, rows, pf, nt)
                    if not c.problems():
                        out.add(c)
        if base.mfma:
            base = dataclasses.replace(base, tm=min(kc, 8), tn=min(mc, 8 if dtype == "f64" else 4))
            if (kc // base.tm) * (mc // base.tn) > WAVE:
                base = dataclasses.replace(base, tn=min(mc, 8))
            rg = THREADS // base.ntiles
            base = dataclasses.replace(base, rows=max(rg, min(base.rows, rg * 8)))
    tiles = {(base.tm, base.tn), (min(kc, 4), min(mc, 4)), (min(kc, 8), min(mc, 8)),
             (min(kc, 4), min(mc, 8)), (min(kc, 8), min(mc, 4))}
    for tm, tn in tiles:
        for rows in (base.rows // 2, base.rows, base.rows * 2):
            for pf in (1, 2):
                # streaming loads for every tile at two tiles in flight, and
                # for the default tile at either depth
                nts = (0, 1) if (tm, tn) == (base.tm, base.tn) or pf == 2 else (0,)
                for nt in nts:
                    c = KernelConfig(dtype, kc, mc, tm, tn, rows, pf, nt)
                    if not c.problems():
                        out.add(c)
            # four tiles in flight (streaming loads) for the register tiles of
            # at most 32 accumulators, whose registers leave room for them
            if tm * tn <= 32 and rows <= base.rows:
                c = KernelConfig(dtype, kc, mc, tm, tn, rows, 4, 1)
                if not c.problems():
                    out.add(c)
    out.update(direct_variants(dtype, kc, mc))
    return sorted(out)


def padded_mfma_variants(dtype: str) -> list[KernelConfig]:
    """Matrix-core kernels of the classes with one side narrower than 16
    (4 and 8 for BF16, 4 and 8 for FP32, 2, 4 and 8 for FP64) and the other
    of 16 or more, the narrow side padded to 16 in LDS.  Their register-tile kernels need up to 64 FMAs per element read
    (plus the BF16 unpacking), more than the vector units sustain at the HBM
    rate for FP32 and BF16.  Three tile sizes around 32 KiB of staged rows,
    one or two tiles in flight, streaming loads."""
    out = []
    # (8-byte rows, two to a 16-byte chunk, are padded too; of those only
    # BF16 at 4 has register tiles that the vector units limit)
    narrow = [w for w in (2, 4, 8)
              if (w * ELEM_BYTES[dtype]) % 16 == 0 or (dtype == "bf16" and w == 4)]
    for kc, mc in ([(n, w) for n in narrow for w in (16, 32, 64)]
                   + [(w, n) for n in narrow for w in (16, 32, 64)]):
        mrows = _pow2_floor(32768 // ((max(kc, 16) + max(mc, 16)) * ELEM_BYTES[dtype]))
        for rows in (mrows // 2, mrows, 2 * mrows):
            for pf in (1, 2):
                c = KernelConfig(dtype, kc, mc, 0, 0, rows, pf, 1)
                if not c.problems():
                    out.append(c)
    return out


def direct_variants(dtype: str, kc: int, mc: int) -> list[KernelConfig]:
    """Direct-engine kernels of a class: every depth with streaming loads (none
    for widths that are not multiples of 16)."""
    rows = 32 if dtype == "bf16" else 4
    depths = DIRECT_DEPTHS_BF16 if dtype == "bf16" else DIRECT_DEPTHS
    out = [KernelConfig(dtype, kc, mc, 0, 0, rows, d, 1, 0, 1) for d in depths]
    return [c for c in out if not c.problems()]


def reference_variants(Ms, tiles, *, dtype: str = "f64") -> list[KernelConfig]:
    """Kernels behind the reference names tsmttsm_{M}_{M}_{TM}_{TM}[...]: basic and
    _sg map to pf=1, _lf and _sg_lf to pf=2 (sycl_codegen_new.py:640-661)."""
    out = []
    for M in Ms:
        for TM in tiles:
            if M % TM:
                continue  # the reference skips TM that does not divide M (:645-647)
            kc = width_class(M)
            base = default_config(dtype, kc, kc)
            for pf in (1, 2):
                # the largest valid row tile not above the default's, else the smallest above
                cands = [KernelConfig(dtype, kc, kc, TM, TM, base.rows >> s, pf) for s in range(8)]
                cands += [KernelConfig(dtype, kc, kc, TM, TM, base.rows << s, pf) for s in range(1, 6)]
                ok = [c for c in cands if c.rows >= 1 and not c.problems()]
                if ok:
                    out.append(ok[0])
    return out


TUNING_FORMAT = 1  # version of the tile-table layout below


def load_tuning(path: str = TUNING_JSON) -> dict:
    """The tile table: {"format", "device", "rocm", "library_version",
    "generated_by", "entries": [A^T B class entries], "sym_entries": [A^T A]}.

    A table of another format, or measured on another architecture than
    ARCH, is rejected: its configurations would be baked into kernels tuned
    for a different chip."""
    if not os.path.exists(path):
        return {"format": TUNING_FORMAT, "device": ARCH, "entries": []}
    with open(path) as f:
        t = json.load(f)
    fmt = t.get("format", 1)
    if fmt != TUNING_FORMAT:
        raise ValueError(f"{path}: tile-table format {fmt}, this generator reads {TUNING_FORMAT}")
    dev = t.get("device", ARCH).split(":")[0]
    if dev != ARCH:
        raise ValueError(f"{path}: measured on {dev}, the library is built for {ARCH}")
    return t


def sym_variants(defaults: dict, dtype: str, kc: int) -> list[KernelConfig]:
    """A^T A variants the autotuner chooses from for a square tuned class: the
    tuned A^T B consumer of the class with half, one, two and four times its
    rows per tile and one or two tiles in flight, and the direct engine's
    kernels of the class (matrix-core widths).  A^T A reads half the bytes of
    A^T B per row, so the direct engine, which keeps every row's operands in
    registers, can win where A^T B prefers the staged kernel.  At 8 the
    staged matrix-core consumer on A's rows padded to 16 is offered too: the
    register tiles there do 8 FMAs per element read (plus BF16 unpacking)."""
    c = defaults[(dtype, kc, kc)][0]
    direct = {dataclasses.replace(v, sym=1) for v in direct_variants(dtype, kc, kc)}
    if c.direct:
        return sorted(direct)
    out = set(direct)
    for rows in (c.rows // 2, c.rows, 2 * c.rows, 4 * c.rows):
        for pf in (1, 2):
            v = dataclasses.replace(c, sym=1, rows=rows, pf=pf)
            if rows >= 1 and not v.problems():
                out.add(v)
    if kc == 8:
        mrows = _pow2_floor(32768 // (16 * ELEM_BYTES[dtype]))
        for rows in (mrows // 2, mrows, 2 * mrows):
            for pf in (1, 2):
                v = KernelConfig(dtype, 8, 8, 0, 0, rows, pf, 1, sym=1)
                if not v.problems():
                    out.add(v)
    return sorted(out)


def sym_kernels(defaults: dict, tuning: dict) -> tuple[list[KernelConfig], dict]:
    """All A^T A kernels of the library and the default per (dtype, class).

    Without a tuned entry the default is the tuned A^T B configuration of the
    square class (same consumer, grid and tile order); a stage holds A only,
    so register-tile kernels take twice the rows where that is valid, which
    keeps the bytes in flight per tile.  Tuned entries come from the
    "sym_entries" of the tile table."""
    cfgs: set[KernelConfig] = set()
    out = {}
    fbs = library_fallbacks(tuning, defaults)
    for dt in DTYPES:
        for kc in CLASSES:
            c, wpc, il = defaults[(dt, kc, kc)]
            s = sym_of(c)
            cfgs.add(s)
            cfgs.add(sym_of(fbs[(dt, kc, kc)][0]))
            out[(dt, kc)] = (s, wpc, il)
            if (kc, kc) in TUNE_SHAPES[dt]:
                cfgs.update(sym_variants(defaults, dt, kc))
    for e in tuning.get("sym_entries", []):
        c = config_of_entry(e, sym=1)
        cfgs.add(c)
        out[(c.dtype, c.kc)] = (c, int(e.get("wgs_per_cu", 1)), int(e.get("interleave", 0)))
    return sorted(cfgs), out


def sym_of(c: KernelConfig) -> KernelConfig:
    """The A^T A kernel of an A^T B configuration: same consumer; a stage
    holds A only, so register-tile kernels take twice the rows where that is
    valid, which keeps the bytes in flight per tile."""
    s = dataclasses.replace(c, sym=1)
    if not c.mfma:
        wide = dataclasses.replace(s, rows=2 * c.rows)
        if not wide.problems():
            s = wide
    if s.problems():
        raise ValueError(f"{s.name}: {s.problems()}")
    return s


def config_of_entry(e: dict, sym: int = 0) -> KernelConfig:
    """KernelConfig of a tile-table entry (or of its "fallback")."""
    c = KernelConfig(e["dtype"], e["kc"], e.get("mc", e["kc"]) if not sym else e["kc"],
                     e["tm"], e["tn"], e["rows"], e["pf"], int(e.get("nt", 0)), sym,
                     int(e.get("direct", 0)))
    if c.problems():
        raise ValueError(f"tuning entry {e} is invalid: {c.problems()}")
    return c


def library_fallbacks(tuning: dict, defaults: dict) -> dict:
    """Per class, the (kernel, interleave) that runs when the inputs do not
    suit the class default: the direct engine reads only dense rows (k and m
    equal to the class widths, leading dimensions k and m, 16-byte aligned
    bases); other kernels handle every input themselves.  A direct default
    falls back to the tuned staged kernel recorded with it, else to the
    heuristic configuration of the class."""
    out = {}
    tuned = {(e["dtype"], e["kc"], e["mc"]): e for e in tuning.get("entries", [])}
    for key, (c, _, il) in defaults.items():
        if not c.direct:
            out[key] = (c, il)
            continue
        fb = tuned.get(key, {}).get("fallback")
        if fb:
            out[key] = (config_of_entry(dict(fb, dtype=key[0], kc=key[1], mc=key[2])),
                        int(fb.get("interleave", 0)))
        else:
            out[key] = (default_config(*key), 0)
    return out


def library_kernels(tuning: dict) -> tuple[list[KernelConfig], dict]:
    """The full A^T B kernel set of the library and the tuned default per class."""
    cfgs: set[KernelConfig] = set()
    defaults: dict[tuple[str, int, int], tuple[KernelConfig, int, int]] = {}
    for dt in DTYPES:
        for kc in CLASSES:
            for mc in CLASSES:
                c = default_config(dt, kc, mc)
                cfgs.add(c)
                defaults[(dt, kc, mc)] = (c, 2, 0)
                # one direct kernel in every class the engine serves; without
                # a tuned entry it is the default from 1024 outputs on (one
                # workgroup per CU), the staged kernel its fallback
                direct = [d for d in direct_variants(dt, kc, mc) if d.pf == (4 if dt == "bf16" else 8)]
                cfgs.update(direct)
                if direct and kc * mc >= 1024 and dt != "bf16":
                    defaults[(dt, kc, mc)] = (direct[0], 1, 0)
        for kc, mc in TUNE_SHAPES[dt]:
            cfgs.update(tuning_variants(dt, kc, mc))
        cfgs.update(padded_mfma_variants(dt))
    cfgs.update(reference_variants([4, 8, 16], [2, 4]))
    for e in tuning.get("entries", []):
        c = config_of_entry(e)
        cfgs.add(c)
        defaults[(c.dtype, c.kc, c.mc)] = (c, int(e.get("wgs_per_cu", 2)),
                                           int(e.get("interleave", 0)))
    cfgs.update(c for c, _ in library_fallbacks(tuning, defaults).values())
    return sorted(cfgs), defaults


# ---------------------------------------------------------------------------
# Emitters
# ---------------------------------------------------------------------------
BANNER = ("// GENERATED by tsm_benchmark_amd/codegen.py - do not edit.\n"
          "// Regenerate with: python -m tsm_benchmark_amd.codegen --library\n")


def emit_kernel(c: KernelConfig) -> str:
    t = CXX_TYPE[c.dtype]
    if c.direct:
        targs = f"{t}, {c.kc}, {c.mc}, {c.pf}, {c.nt}" + (", 1" if c.sym else "")
        return (
            f"// {c.name}: KC={c.kc} MC={c.mc}, matrix cores fed from global memory, "
            f"{c.pf} 4-row steps in flight per wave, nt={c.nt}, LDS {c.lds_bytes} B\n"
            f"extern \"C\" __global__ void __launch_bounds__(256) {c.name}(tsm::AtbArgs a) {{\n"
            f"  tsm::direct_body<{targs}>(a);\n"
            f"}}\n"
            f"hipError_t launch_{c.name}(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {{\n"
            f"  hipLaunchKernelGGL({c.name}, dim3(grid), dim3(256), 0, s, a);\n"
            f"  return hipGetLastError();\n"
            f"}}\n")
    targs = f"{t}, {c.kc}, {c.mc}, {c.tm}, {c.tn}, {c.rows}, {c.pf}, {c.nt}" + (", 1" if c.sym else "")
    if c.mfma:
        shape = (f"//   matrix cores: {c.rows // c.kstep} K-steps of {c.kstep} rows per tile, "
                 f"LDS {c.lds_bytes} B\n")
    else:
        shape = (f"//   {c.ntiles} register tiles x {c.row_lanes} row lanes, "
                 f"{c.rows // c.row_lanes} rows/thread/tile, LDS {c.lds_bytes} B\n")
    return (
        f"// {c.name}: KC={c.kc} MC={c.mc} TM={c.tm} TN={c.tn} rows={c.rows} pf={c.pf} nt={c.nt}\n"
        + shape +
        f"extern \"C\" __global__ void __launch_bounds__(256) {c.name}(tsm::AtbArgs a) {{\n"
        f"  tsm::atb_body<{targs}>(a);\n"
        f"}}\n"
        f"hipError_t launch_{c.name}(const tsm::AtbArgs& a, unsigned grid, hipStream_t s) {{\n"
        f"  hipLaunchKernelGGL({c.name}, dim3(grid), dim3(256), 0, s, a);\n"
        f"  return hipGetLastError();\n"
        f"}}\n")


def emit_kernel_file(dtype: str, kc: int, mc: int, cfgs: list[KernelConfig]) -> str:
    body = [BANNER, '#include "../tsm_internal.hpp"\n']
    for c in cfgs:
        if c.dtype == dtype and c.kc == kc and c.mc == mc and not c.direct:
            body.append(emit_kernel(c))
    return "\n".join(body)


# Direct-engine kernels are compiled with MFMA results in ordinary VGPRs
# (build.py adds -mllvm -amdgpu-mfma-vgpr-form=true for these files): with
# the default AGPR form the accumulators are copied between the two register
# files on every loop iteration.
DIRECT_FILE = "tsm_direct_{}.hip"


def emit_direct_file(dtype: str, cfgs: list[KernelConfig]) -> str:
    body = [BANNER, '#include "../tsm_internal.hpp"\n']
    for c in cfgs:
        if c.dtype == dtype and c.direct:
            body.append(emit_kernel(c))
    return "\n".join(body)


# The batched kernel carries every class's registers and runs one workgroup
# per CU, so LDS does not bound its occupancy: its FP64 and FP32 classes take
# row tiles of up to BATCH_STAGE bytes (the bytes in flight per CU that a
# second workgroup would otherwise add).
BATCH_LDS = 81920
BATCH_STAGE = 40960


def batched_classes(defaults: dict, dtype: str, tuning: dict | None = None) -> list[KernelConfig]:
    """Per width class, the configuration the batched kernel uses: the tuned
    default (its staged fallback for a direct-engine default, whose lanes
    each stream their own rows) when it fits the batched LDS budget, else the
    heuristic one."""
    fbs = library_fallbacks(tuning or {}, defaults)
    out = []
    # FP32/BF16 carry a (hi, lo) pair per accumulator: their 64-wide classes
    # would not fit the shared register budget, so batching stops at 32.
    widths = CLASSES if dtype == "f64" else tuple(c for c in CLASSES if c <= 32)
    for kc in widths:
        for mc in widths:
            c = fbs[(dtype, kc, mc)][0]
            # (the padded matrix-core kernels of the 8-wide sides stay out:
            # their zeroed LDS padding would be refilled for every item)
            if (c.lds_bytes > BATCH_LDS or (dtype != "f64" and c.tm * c.tn > 32)
                    or (c.mfma and min(kc, mc) < 16)):
                c = default_config(dtype, kc, mc)
            while dtype != "bf16":  # BF16's wider class bodies would spill
                wider = dataclasses.replace(c, rows=2 * c.rows)
                if (wider.stage_bytes > BATCH_STAGE or wider.lds_bytes > BATCH_LDS
                        or wider.problems()):
                    break
                c = wider
            out.append(c)
    return out


def emit_batched_file(dtype: str, classes: list[KernelConfig]) -> str:
    t = CXX_TYPE[dtype]
    lds = max(c.lds_bytes for c in classes)
    name = f"tsm_atb_batched_{dtype}"
    out = [BANNER, '#include "../tsm_internal.hpp"\n',
           f"// {name}: {len(classes)} width classes, LDS {lds} B",
           f"extern \"C\" __global__ void __launch_bounds__(256) {name}(tsm::BatchArgs ba) {{",
           f"  __shared__ __attribute__((aligned(16))) unsigned char smem[{lds}];",
           "  __shared__ int s_flag;",
           "  __shared__ unsigned s_next;",
           "  for (int64_t i = blockIdx.x; i < ba.n_items; i = (int64_t)gridDim.x + s_next) {",
           "    const tsm::BatchItem it = ba.items[i];",
           "    switch (ba.probs[it.problem].cls) {"]
    for i, c in enumerate(classes):
        out.append(f"      case {i}: tsm::batched_item_call<{t}, {c.kc}, {c.mc}, {c.tm}, {c.tn}, {c.rows}, "
                   f"{c.pf}, {c.nt}>(ba, it, smem, &s_flag, &s_next); break;  // {c.name}")
    out += ["      default: tsm::batched_skip(ba, &s_next); break;", "    }", "  }",
            "  tsm::batched_exit(ba);", "}", "",
            f"static const tsm::BatchClass kClasses[] = {{"]
    out += [f"  {{{c.kc}, {c.mc}, {c.rows}}}," for c in classes]
    out += ["};", "",
            f"hipError_t launch_{name}(const tsm::BatchArgs& a, unsigned grid, hipStream_t s) {{",
            f"  hipLaunchKernelGGL({name}, dim3(grid), dim3(256), 0, s, a);",
            "  return hipGetLastError();", "}", "",
            f"const tsm::BatchClass* classes_{name}(int* count) {{",
            f"  *count = {len(classes)};", "  return kClasses;", "}", "",
            f"int occupancy_{name}() {{",
            "  int nb = 0;",
            f"  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>({name}), 256, 0) !=",
            "          hipSuccess || nb < 1)",
            "    nb = 1;",
            "  return nb;", "}", ""]
    return "\n".join(out)


def emit_table(cfgs: list[KernelConfig], defaults: dict, tuning: dict) -> str:
    idx = {c: i for i, c in enumerate(cfgs)}
    sym_cfgs, sym_defaults = sym_kernels(defaults, tuning)
    sidx = {c: i for i, c in enumerate(sym_cfgs)}
    fbs = library_fallbacks(tuning, defaults)
    out = [BANNER, '#include "../tsm_internal.hpp"\n']
    for c in cfgs + sym_cfgs:
        out.append(f"hipError_t launch_{c.name}(const tsm::AtbArgs&, unsigned, hipStream_t);")

    def entry(c: KernelConfig) -> str:
        return (f"  {{\"{c.name}\", {DTYPE_ID[c.dtype]}, {c.kc}, {c.mc}, {c.tm}, {c.tn}, "
                f"{c.rows}, {c.pf}, {c.nt}, {c.lds_bytes}, launch_{c.name}, {c.engine}}},")
    out.append("\nstatic const tsm::KernelEntry kTable[] = {")
    out += [entry(c) for c in cfgs]
    out.append("};\n")
    out.append("static const tsm::KernelEntry kSym[] = {")
    out += [entry(c) for c in sym_cfgs]
    out.append("};\n")
    out.append("static const tsm::TunedEntry kSymTuned[] = {")
    for (dt, kc), (c, wpc, il) in sorted(sym_defaults.items()):
        fc, fil = (sym_of(fbs[(dt, kc, kc)][0]), fbs[(dt, kc, kc)][1]) if c.direct else (c, il)
        out.append(f"  {{{DTYPE_ID[dt]}, {kc}, {kc}, {sidx[c]}, {wpc}, {il}, {sidx[fc]}, {fil}}},"
                   f"  // {c.name}" + (f" (else {fc.name})" if fc != c else ""))
    out.append("};\n")
    out.append("static const tsm::TunedEntry kTuned[] = {")
    for (dt, kc, mc), (c, wpc, il) in sorted(defaults.items()):
        fc, fil = fbs[(dt, kc, mc)]
        out.append(f"  {{{DTYPE_ID[dt]}, {kc}, {mc}, {idx[c]}, {wpc}, {il}, {idx[fc]}, {fil}}},"
                   f"  // {c.name}" + (f" (else {fc.name})" if fc != c else ""))
    out.append("};\n")
    for dt in DTYPES:
        nm = f"tsm_atb_batched_{dt}"
        out.append(f"hipError_t launch_{nm}(const tsm::BatchArgs&, unsigned, hipStream_t);")
        out.append(f"const tsm::BatchClass* classes_{nm}(int* count);")
        out.append(f"int occupancy_{nm}();")
    out.append("static const tsm::BatchedKernel* batched_all() {")
    out.append("  static tsm::BatchedKernel all[3];")
    out.append("  static bool init = [] {")
    for i, dt in enumerate(DTYPES):
        nm = f"tsm_atb_batched_{dt}"
        lds = max(c.lds_bytes for c in batched_classes(defaults, dt, tuning))
        out.append(f"    all[{i}].name = \"{nm}\"; all[{i}].dtype = {DTYPE_ID[dt]};")
        out.append(f"    all[{i}].classes = classes_{nm}(&all[{i}].n_classes);")
        out.append(f"    all[{i}].lds_bytes = {lds}; all[{i}].launch = launch_{nm};")
        out.append(f"    all[{i}].occupancy = occupancy_{nm};")
    out.append("    return true;")
    out.append("  }();")
    out.append("  (void)init;")
    out.append("  return all;")
    out.append("}")
    out.append("extern \"C\" const tsm::BatchedKernel* tsm_gen_batched_table(int* count) {\n"
               f"  *count = {len(DTYPES)};\n  return batched_all();\n}}")
    out.append("extern \"C\" const tsm::KernelEntry* tsm_gen_kernel_table(int* count) {\n"
               "  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));\n  return kTable;\n}")
    out.append("extern \"C\" const tsm::KernelEntry* tsm_gen_sym_table(int* count) {\n"
               "  *count = (int)(sizeof(kSym) / sizeof(kSym[0]));\n  return kSym;\n}")
    out.append("extern \"C\" const tsm::TunedEntry* tsm_gen_sym_tuned_table(int* count) {\n"
               "  *count = (int)(sizeof(kSymTuned) / sizeof(kSymTuned[0]));\n  return kSymTuned;\n}")
    out.append("extern \"C\" const tsm::TunedEntry* tsm_gen_tuned_table(int* count) {\n"
               "  *count = (int)(sizeof(kTuned) / sizeof(kTuned[0]));\n  return kTuned;\n}")
    return "\n".join(out) + "\n"


def _write_if_changed(path: str, text: str) -> bool:
    if os.path.exists(path):
        with open(path) as f:
            if f.read() == text:
                return False
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)
    return True


def generate_library(out_dir: str = GEN_DIR, tuning_path: str = TUNING_JSON) -> list[str]:
    """Write the library's generated sources; returns the .hip/.cpp paths."""
    tuning = load_tuning(tuning_path)
    cfgs, defaults = library_kernels(tuning)
    syms, _ = sym_kernels(defaults, tuning)
    paths = []
    for dt in DTYPES:
        for kc in CLASSES:
            for mc in CLASSES:
                p = os.path.join(out_dir, f"tsm_kernels_{dt}_{kc}x{mc}.hip")
                _write_if_changed(p, emit_kernel_file(dt, kc, mc, cfgs + syms))
                paths.append(p)
    for dt in DTYPES:
        p = os.path.join(out_dir, DIRECT_FILE.format(dt))
        _write_if_changed(p, emit_direct_file(dt, cfgs + syms))
        paths.append(p)
    keep = {os.path.basename(x) for x in paths} | {"tsm_table.hip"} | \
        {f"tsm_batched_{dt}.hip" for dt in DTYPES}
    for f in os.listdir(out_dir):  # drop files of an older split
        if f.endswith(".hip") and f not in keep:
            os.remove(os.path.join(out_dir, f))
    for dt in DTYPES:
        p = os.path.join(out_dir, f"tsm_batched_{dt}.hip")
        _write_if_changed(p, emit_batched_file(dt, batched_classes(defaults, dt, tuning)))
        paths.append(p)
    p = os.path.join(out_dir, "tsm_table.hip")
    _write_if_changed(p, emit_table(cfgs, defaults, tuning))
    paths.append(p)
    return paths


# ---- name-compatible layer (the reference's generated header role) ----------
def emit_name_layer(Ms, tiles, stride, unroll, with_benchmark: bool) -> str:
    """The reference's kernel names as inline wrappers over the tuned launch.

    Each wrapper keeps the reference's signature shape ``(queue, K, A, B, C)``
    and contract ``C += A^T B`` (sycl_codegen_new.py:158-159, 218-225); the queue
    becomes a stream and the returned value a tsm_status.  The register tile,
    leap-frog depth, stride and unroll of each name are recorded in
    ``kGeneratedKernels``; the launch is the library's tuned one."""
    lines = ["#pragma once", "// GENERATED by tsm_benchmark_amd/codegen.py --name-layer",
             f"// M={list(Ms)} tile-sizes={list(tiles)} stride={stride} unroll={unroll}",
             "#include \"tsm/tsmttsm.hh\"", "", "namespace tsm_compat {", ""]
    names = []
    for M in Ms:
        for TM in tiles:
            if M % TM:
                continue
            for suffix, lf, sg in (("", 0, 0), ("_lf", 1, 0), ("_sg", 0, 1), ("_sg_lf", 1, 1)):
                name = f"tsmttsm_{M}_{M}_{TM}_{TM}{suffix}"
                names.append((name, M, TM, lf, sg))
                lines += [
                    "template <typename T>",
                    f"inline tsm_status {name}(tsm_stream_t q, int64_t K, const T* A, const T* B, T* C) {{",
                    f"  return tsm_compat::generated_call<T>(q, K, {M}, {M}, A, B, C);",
                    "}", ""]
    if with_benchmark:
        lines += ["// Every generated name with its FP64 instance, in generation order (the",
                  "// generated benchmark's list, sycl_codegen_new.py:478-487).",
                  "// The reference's decomposition of each name is recorded (tile TM x TM,",
                  "// leap-frog prefetch, sub-group reduction, stride, unroll).",
                  "struct GeneratedKernel {",
                  "  const char* name;",
                  "  int M, TM, leap_frog, sub_group;",
                  "  long long stride;",
                  "  int unroll;",
                  "  tsm_status (*fn)(tsm_stream_t, int64_t, const double*, const double*, double*);",
                  "};",
                  "inline const GeneratedKernel kGeneratedKernels[] = {"]
        lines += [f"  {{\"{n}\", {M}, {TM}, {lf}, {sg}, {stride}, {unroll}, &{n}<double>}},"
                  for n, M, TM, lf, sg in names]
        lines += ["};", "inline const char* const kGeneratedNames[] = {"]
        lines += [f"  \"{n}\"," for n, *_ in names]
        lines += ["};", ""]
    lines += ["}  // namespace tsm_compat", ""]
    return "\n".join(lines)


def suggest_stride(Ms, tiles, local_size=THREADS, sg_size=WAVE, target_wgs=2048) -> int:
    """The reference's stride rule (sycl_codegen_new.py:563-580): the smallest
    multiple of lcm(local, num_tiles, sg*num_tiles) near local*target_wgs.  On
    gfx950 the stride is the persistent grid in work-items and only sets
    ``tsm_kernel_config.grid``."""
    req = local_size
    for M in Ms:
        for TM in tiles:
            if M % TM:
                continue
            nt = (M // TM) ** 2
            req = lcm(req, nt)
            req = lcm(req, sg_size * nt)
    full = lcm(req, local_size)
    n = max(1, (local_size * target_wgs + full - 1) // full)
    return n * full


def validate_stride(Ms, tiles, stride, local_size=THREADS, sg_size=WAVE) -> list[str]:
    """validate_stride (sycl_codegen_new.py:530-560); returns the error lines."""
    errs = []
    if stride % local_size:
        return [f"error: the stride {stride} is not a whole number of work-groups of "
                f"{local_size} items"]
    for M in Ms:
        for TM in tiles:
            if M % TM:
                continue
            nt = (M // TM) ** 2
            if stride % nt:
                errs.append(f"note: M={M}, TM={TM} has {nt} register tiles, which do not "
                            f"split the stride {stride} evenly")
            if (stride // sg_size) % nt:
                errs.append(f"note: M={M}, TM={TM} has {nt} register tiles, which do not "
                            f"split the {stride // sg_size} sub-groups of the stride evenly")
    return errs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Generate gfx950 HIP kernels for C = A^T B")
    ap.add_argument("--library", action="store_true",
                    help="write the library kernel set into csrc/generated/")
    ap.add_argument("--tuning", default=TUNING_JSON, help="tile-table JSON (autotuner output)")
    ap.add_argument("--out-dir", default=GEN_DIR)
    ap.add_argument("--name-layer", metavar="FILE",
                    help="write the reference-name wrapper header to FILE ('-' = stdout)")
    ap.add_argument("--M", type=int, nargs="+", default=[4, 8, 16])
    ap.add_argument("--tile-sizes", type=int, nargs="+", default=[2, 4])
    ap.add_argument("--stride", type=int, default=None)
    ap.add_argument("--local-size", type=int, default=THREADS)
    ap.add_argument("--sg-size", type=int, default=WAVE)
    ap.add_argument("--unroll", type=int, default=1)
    ap.add_argument("--with-benchmark", action="store_true")
    ap.add_argument("--list", action="store_true", help="print the library kernel names")
    args = ap.parse_args(argv)

    if args.local_size != THREADS:
        print(f"Error: --local-size must be {THREADS} on gfx950", file=sys.stderr)
        return 1
    if args.sg_size != WAVE:
        print(f"Error: --sg-size must be {WAVE} (gfx950 wavefront)", file=sys.stderr)
        return 1
    if args.local_size % args.sg_size:
        print("Error: --local-size must be a multiple of --sg-size", file=sys.stderr)
        return 1
    if args.unroll < 1:
        print("Error: --unroll must be >= 1", file=sys.stderr)
        return 1
    for M in args.M:
        width_class(M)
    if args.stride is None:
        args.stride = suggest_stride(args.M, args.tile_sizes, args.local_size, args.sg_size)
    errs = validate_stride(args.M, args.tile_sizes, args.stride, args.local_size, args.sg_size)
    if errs:
        for e in errs:
            print(e, file=sys.stderr)
        return 1

    if args.library:
        for p in generate_library(args.out_dir, args.tuning):
            print(p)
    if args.name_layer:
        text = emit_name_layer(args.M, args.tile_sizes, args.stride, args.unroll,
                               args.with_benchmark)
        if args.name_layer == "-":
            sys.stdout.write(text)
        else:
            _write_if_changed(args.name_layer, text)
    if args.list:
        cfgs, _ = library_kernels(load_tuning(args.tuning))
        for c in cfgs:
            print(c.name)
    return 0


if __name__ == "__main__":
    sys.exit(main())
