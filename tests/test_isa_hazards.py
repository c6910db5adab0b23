"""SGPR wait states of the hand-written carry chains, and scratch use, checked
in the gfx950 code.

The Montgomery products in natural_number_amd/csrc/nnh_arith.hpp are inline
assembly that carries 64-bit column sums through SGPR carry masks
(v_mad_u64_u32 / v_add_co / v_addc / v_subb with an SGPR carry-out read by a
later instruction).  On gfx950 a VALU instruction that reads an SGPR written
by another VALU instruction needs two wait states between them; the compiler
inserts them (s_nop) for the code it generates, but not inside inline
assembly, whose blocks the source orders and pads by hand.

These tests compile every HIP translation unit of the library to gfx950
assembly (device code only, the library's flags) and scan each
straight-line run of instructions: every VALU read of an SGPR (or VCC) that a
VALU instruction wrote must come at least two wait states after the write
(each instruction counts one, s_nop N counts N + 1).  It fails if a compiler
or scheduling change ever places the blocks so that a carry is read too early.
A sensitivity check asks the same scan for three wait states and expects it
to find the carry reads of the inline blocks that sit exactly at two, so that
a scanner which parsed nothing cannot pass.  The same assembly's kernel
metadata must show no VGPR or SGPR spills and no private (scratch) segment
for any kernel, and no transform or CRT kernel may access memory with flat
instructions.  No GPU is needed.
"""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "natural_number_amd", "csrc")
# units whose kernels run the inline carry chains, and every other HIP unit
ASM_UNITS = ["nnh_mid_lazy.hip", "nnh_mid.hip", "nnh_top_lazy.hip", "nnh_top.hip", "nnh_crt.hip", "nnh_peak.hip"]
UNITS = ASM_UNITS + ["nnh_nat.hip", "nnh_mlp.hip"]
FLAGS = ["-O3", "-std=c++20", "--offload-arch=gfx950", "--cuda-device-only", "-S",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# VOP3 forms whose second operand is an SGPR destination (carry-out)
_SDST = re.compile(r"^v_(add|sub|subrev|addc|subb|subbrev)_co_u32_e64$|^v_mad_[ui]64_[ui]32$|^v_div_scale")
_BRANCH = ("s_cbranch", "s_branch", "s_setpc_b64", "s_endpgm")


def _sgprs(tok):
    tok = tok.strip()
    m = re.match(r"^s\[(\d+):(\d+)\]$", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"^s(\d+)$", tok)
    if m:
        return {int(m.group(1))}
    if tok == "vcc":
        return {"vcc_lo", "vcc_hi"}
    if tok in ("vcc_lo", "vcc_hi"):
        return {tok}
    return set()


def scan(text, need=2):
    """Return (violations, inline_reads_at): each violation is (function,
    line, wait states, instruction); inline_reads_at counts VALU SGPR reads
    inside inline-asm blocks by their wait-state distance from the write."""
    viol, dist = [], {}
    func, last_w, pos, in_asm = None, {}, 0, False
    for ln, raw in enumerate(text.splitlines(), 1):
        st = raw.strip()
        if st.startswith(";;#ASMSTART"):
            in_asm = True
            continue
        if st.startswith(";;#ASMEND"):
            in_asm = False
            continue
        s = raw.split(";")[0].strip()
        if not s:
            continue
        if s.endswith(":"):
            # a function starts afresh; a block label keeps the state of the
            # code that falls through to it (a branch resets it, below)
            if not s.startswith("."):
                func, last_w = s[:-1], {}
            continue
        if s.startswith("."):
            continue
        parts = s.split(None, 1)
        mn = parts[0]
        ops = [o.strip() for o in parts[1].split(",")] if len(parts) > 1 else []
        if mn == "s_nop":
            pos += int(ops[0], 0) + 1
            continue
        pos += 1
        if mn.startswith(_BRANCH):
            last_w = {}
            continue
        if not mn.startswith("v_"):
            if ops:  # a scalar write replaces the VALU value of that register
                for r in _sgprs(ops[0]):
                    last_w.pop(r, None)
            continue
        dst, src = set(), ops[1:]
        if mn.startswith("v_cmp"):
            dst = _sgprs(ops[0]) if ops else set()
        elif _SDST.match(mn):
            dst, src = _sgprs(ops[1]), ops[2:]
        elif mn.endswith("_e32") and "_co_" in mn and len(ops) > 1 and ops[1] == "vcc":
            dst, src = _sgprs("vcc"), ops[2:]
        elif mn.startswith(("v_readfirstlane", "v_readlane")):
            dst = _sgprs(ops[0])
        if mn.endswith("_e32") and mn.startswith(("v_cndmask", "v_addc", "v_subb")):
            src = src + ["vcc"]
        near = [pos - last_w[r] - 1 for o in src for r in _sgprs(o) if r in last_w]
        if near:
            ws = min(near)
            if in_asm:
                dist[ws] = dist.get(ws, 0) + 1
            if ws < need:
                viol.append((func, ln, ws, s))
        for r in dst:
            last_w[r] = pos
    return viol, dist


def _compile(src, out_dir):
    out = os.path.join(out_dir, os.path.splitext(src)[0] + ".s")
    r = subprocess.run(["hipcc", src, "-o", out] + FLAGS, cwd=CSRC, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return src, open(out).read()


@pytest.fixture(scope="module")
def gfx950_asm(tmp_path_factory):
    out_dir = str(tmp_path_factory.mktemp("isa"))
    jobs = max(1, min(len(UNITS), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        return dict(ex.map(lambda u: _compile(u, out_dir), UNITS))


def test_inline_carry_chains_keep_two_wait_states(gfx950_asm):
    bad = {}
    for unit, text in gfx950_asm.items():
        assert unit not in ASM_UNITS or ";;#ASMSTART" in text, f"{unit}: no inline assembly found"
        viol, _ = scan(text, need=2)
        if viol:
            bad[unit] = viol[:10]
    assert not bad, f"VALU SGPR reads closer than two wait states to their write: {bad}"


def test_scan_sees_the_inline_carries(gfx950_asm):
    # the carry chains are padded to exactly two wait states in places: a
    # scan that asks for three must report those reads
    text = gfx950_asm["nnh_mid_lazy.hip"]
    viol, dist = scan(text, need=3)
    assert dist.get(2, 0) > 0, dist
    assert viol


def test_scan_flags_a_planted_early_read():
    # a carry read right after its write, as a block without its s_nop would be
    text = "\n".join([
        "k:",
        ";;#ASMSTART",
        "v_add_co_u32_e64 v1, s[4:5], v2, v3",
        "v_addc_co_u32_e64 v6, s[8:9], v7, 0, s[4:5]",
        ";;#ASMEND",
    ])
    viol, _ = scan(text, need=2)
    assert len(viol) == 1 and viol[0][2] == 0
    padded = text.replace("v_addc", "s_nop 1\nv_addc")
    assert scan(padded, need=2)[0] == []


def spills(text):
    """(kernel, VGPR spills, SGPR spills, private segment bytes) for every
    kernel of a gfx950 assembly file whose code uses scratch memory."""
    meta = text[text.index("amdhsa.kernels:"):]
    out = []
    for m in re.finditer(r"\.name:\s+(\S+)\n(.*?)\.vgpr_spill_count:\s+(\d+)", meta, re.S):
        body = m.group(2)
        priv = re.search(r"\.private_segment_fixed_size:\s+(\d+)", body)
        ssp = re.search(r"\.sgpr_spill_count:\s+(\d+)", body)
        v, p_, s_ = int(m.group(3)), int(priv.group(1)) if priv else 0, int(ssp.group(1)) if ssp else 0
        if v or p_ or s_:
            out.append((m.group(1), v, s_, p_))
    return out


def test_no_kernel_uses_scratch(gfx950_asm):
    # every kernel keeps its working set in registers and LDS: a spill to
    # scratch memory (the occupancy-capped limb-loading pass once spilled two
    # VGPRs) shows up here before it costs a GPU run
    kernels = 0
    bad = {}
    for unit, text in gfx950_asm.items():
        kernels += text.count(".vgpr_spill_count:")
        sp = spills(text)
        if sp:
            bad[unit] = sp[:10]
    assert kernels > 100, kernels
    assert not bad, f"kernels using scratch memory: {bad}"


def test_spill_scan_reads_the_metadata():
    meta = "\n".join([
        "amdhsa.kernels:",
        "  - .name:           k_a",
        "    .private_segment_fixed_size: 0",
        "    .sgpr_spill_count: 0",
        "    .vgpr_spill_count: 0",
        "  - .name:           k_b",
        "    .private_segment_fixed_size: 12",
        "    .sgpr_spill_count: 0",
        "    .vgpr_spill_count: 2",
    ])
    assert spills(meta) == [("k_b", 2, 0, 12)]


def test_transform_and_crt_kernels_use_global_memory_instructions(gfx950_asm):
    # Pointers read from a job table are generic to the compiler; without the
    # address-space casts of the residue, limb and output accesses the
    # batched kernels fall back to flat instructions, which also count
    # against the LDS wait counter (each LDS transpose then waits for the
    # tile's memory operations).  No transform or CRT kernel may use them.
    bad = {}
    for unit in ("nnh_mid_lazy.hip", "nnh_mid.hip", "nnh_top_lazy.hip", "nnh_top.hip", "nnh_crt.hip"):
        text = gfx950_asm[unit]
        for m in re.finditer(r"^(_ZN\w+):", text, re.M):
            end = text.find(".Lfunc_end", m.end())
            n = len(re.findall(r"^\s+flat_(load|store|atomic)", text[m.end():end], re.M))
            if n:
                bad.setdefault(unit, []).append((m.group(1)[:60], n))
    assert not bad, f"kernels with flat memory instructions: {bad}"
