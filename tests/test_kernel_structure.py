"""Static checks of the gfx950 code the compiler emits for csrc/*.hip (CPU only).

A __syncthreads() is only well defined when every wave of the workgroup reaches
it under uniform control flow. The compiler decides uniformity from what it can
prove: a loop whose exit depends on a value loaded from memory is divergent to
it, even when every lane holds the same value, and it then lowers the loop with
per-lane exec masking around the barrier. That has produced wrong results (the
FP64 stage kernels of a kernel with an empty pair function dropped the LDS
stores of the lanes active in a loop's last partial tile). The kernels route such
values through readfirstlane; this test keeps it that way by flagging any loop
that exits through exec masking (s_andn2_b64 exec, exec, ...) and contains an
s_barrier, in any kernel of this library.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "3bem_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def divergent_barrier_loops(asm_text):
    """(function, loop header) pairs: loops with a divergent exit that contain a barrier.

    Uses the block annotations LLVM writes into the assembly
    ("; in Loop: Header=BB.. Depth=d", "; Parent Loop BB.. Depth=d")."""
    out = []
    fn, blocks, parent = None, {}, {}

    def finish():
        if fn is None:
            return

        def inside(h, top):
            while h:
                if h == top:
                    return True
                h = parent.get(h)
            return False

        divergent = {h for h, lines in blocks.values() if h and any("s_andn2_b64 exec, exec" in l for l in lines)}
        for top in divergent:
            if any(h and inside(h, top) and any("s_barrier" in l for l in lines) for h, lines in blocks.values()):
                out.append((fn, top))

    cur = None
    for line in asm_text.splitlines():
        m = re.match(r"^(_Z\w+):", line)
        if m:
            finish()
            fn, blocks, parent, cur = m.group(1), {"entry": (None, [])}, {}, "entry"
            continue
        # labelled blocks (.LBB..) and unlabelled fall-through blocks ("; %bb.N:"),
        # which carry their own loop annotations
        m = re.match(r"^(?:\.(LBB\w+)|; %(bb\.\d+)):(.*)$", line)
        if m and fn:
            cur, info = m.group(1) or m.group(2), m.group(3)
            h = None
            mh = re.search(r"Header=(BB\w+) Depth=", info)
            if mh:
                h = mh.group(1)
            if "Loop Header" in info or "Parent Loop" in info:
                h = cur[1:]
                mp = re.search(r"Parent Loop (BB\w+) Depth=", info)
                if mp:
                    parent[h] = mp.group(1)
            blocks[cur] = (h, [])
            continue
        if fn and cur:
            blocks[cur][1].append(line)
    finish()
    return out


SYNTHETIC_BAD = """
_Z3badv:
\ts_mov_b32 s0, 0
.LBB0_1:                                ; =>This Loop Header: Depth=1
\ts_barrier
\tv_cmp_lt_i32_e32 vcc, v1, v2
\ts_andn2_b64 exec, exec, vcc
\ts_cbranch_execz .LBB0_2
.LBB0_2:
\ts_endpgm
_Z5afterv:
.LBB2_1:                                ; =>This Inner Loop Header: Depth=1
\tv_add_u32_e32 v1, 1, v1
\ts_andn2_b64 exec, exec, vcc
\ts_cbranch_execnz .LBB2_1
; %bb.2:
\ts_or_b64 exec, exec, s[0:1]
\ts_barrier
\ts_endpgm
_Z4goodv:
.LBB1_1:                                ; =>This Loop Header: Depth=1
\ts_barrier
\ts_cmp_lt_i32 s1, s2
\ts_cbranch_scc1 .LBB1_1
.LBB1_2:                                ; =>This Loop Header: Depth=1
\tv_add_u32_e32 v1, 1, v1
\ts_andn2_b64 exec, exec, vcc
\ts_cbranch_execz .LBB1_3
.LBB1_3:
\ts_endpgm
"""


def test_checker_flags_a_barrier_in_a_divergent_loop():
    # a barrier in a loop with a divergent exit is flagged; a barrier in a scalar
    # loop, or a divergent loop without a barrier, is not
    assert divergent_barrier_loops(SYNTHETIC_BAD) == [("_Z3badv", "BB0_1")]


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_barrier_inside_divergent_loops():
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))

    def asm(src):
        cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "--cuda-device-only", "-S",
               os.path.join(CSRC, src), "-o", "-"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return src, r.stdout

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        results = list(ex.map(asm, srcs))
    bad = []
    for src, text in results:
        assert "s_barrier" in text or src != "stages.hip"
        # this library's kernels (namespace tbem); rocPRIM's sort kernels are out of scope
        bad += [(src, f, h) for f, h in divergent_barrier_loops(text) if f.startswith("_ZN4tbem")]
    assert not bad, bad
