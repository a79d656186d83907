"""Compiled loop structure of the kernels whose blocks loop over work with
block barriers inside (CPU only: hipcc cross-compiles gfx950 assembly).

A `threadIdx.x == 0` region on each side of such a loop's back edge can be
merged by the compiler into a lane-0 latch of its own; structurisation then
turns the loop into nested loops with lane 0 of wave 0 in the outer one and
all other lanes in an inner loop that starts at the barrier.  The other lanes
then pass the barrier without lane 0's update and spin on stale shared state
-- a stall seen on MI355X with an earlier batch kernel.  The kernels now keep
every branch at the edges of these loops wave-uniform.  This test checks the
assembly for that: the block loop's first barrier sits at loop depth 1, and
the loop header branches on a scalar (wave-uniform) condition before it
touches the exec mask."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cuts_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not available")


def assembly(src, tmp_path):
    out = tmp_path / (os.path.basename(src) + ".s")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", src, "-o", str(out)],
                   check=True, capture_output=True, cwd=CSRC)
    return out.read_text()


def function_body(asm, name):
    start = asm.index(f"\n{name}:")
    end = asm.index(".Lfunc_end", start)
    return asm[start:end].splitlines()


def blocks(lines):
    """[label, loop depth, enclosing depth-1 loop header, instructions] per
    basic block, from the loop annotations LLVM writes next to block labels
    (labels are normalised to the `BB<f>_<n>` form those annotations use)."""
    out, cur = [], None
    for ln in lines:
        m = re.match(r"^(\.L(BB\w+)|; %bb\.\d+):(.*)$", ln)
        if m:
            cur = [m.group(2) or m.group(1), 0, None, []]
            out.append(cur)
            note = m.group(3)
        elif cur is not None and ln.strip().startswith(";"):
            note = ln
        else:
            note = ""
            if cur is not None and ln.strip() and not ln.strip().startswith("."):
                cur[3].append(ln.strip())
        if cur is None or cur[1]:
            continue
        if "=>This Loop Header: Depth=1" in note:
            cur[1], cur[2] = 1, cur[0]
            continue
        h = re.search(r"(?:in Loop: Header=(BB\w+)|Loop Header:) Depth=(\d+)", note)
        if h:
            cur[1] = int(h.group(2))
            if cur[1] == 1:
                cur[2] = h.group(1)
    # deeper blocks: the depth-1 loop laid out around them
    header = None
    for b in out:
        if b[1] == 1:
            header = b[2]
        elif b[1] > 1:
            b[2] = header
    return out


def check_block_loop(lines, what):
    bl = blocks(lines)
    first = next((b for b in bl for ins in b[3] if ins.startswith("s_barrier")), None)
    assert first is not None, f"{what}: no barrier found"
    assert first[1] == 1, f"{what}: the block loop's first barrier is at loop depth {first[1]}"
    header = next(b for b in bl if b[0] == first[2])
    for ins in header[3]:
        if ins.startswith("s_cbranch"):
            assert ins.startswith(("s_cbranch_scc", "s_cbranch_vcc")), f"{what}: loop header branches on {ins}"
            break
        assert "exec" not in ins, f"{what}: loop header changes the exec mask before branching ({ins})"
    else:
        raise AssertionError(f"{what}: loop header {header[0]} has no branch")


@pytest.mark.parametrize("width", [7, 11, 15])
def test_batch_level_loop_is_wave_uniform(tmp_path, width):
    asm = assembly(os.path.join(CSRC, "batch.hip"), tmp_path)
    name = f"_ZN4cuts13k_batch_levelILi{width}EEEvNS_9BatchArgsEj"
    check_block_loop(function_body(asm, name), f"k_batch_level<{width}>")


def test_expand_tile_loops_are_wave_uniform(tmp_path):
    asm = assembly(os.path.join(CSRC, "match.hip"), tmp_path)
    for L, mode in ((4, 2), (3, 0), (5, 1), (3, 3), (3, 4), (2, 5), (2, 6)):
        name = f"_ZN4cuts12_GLOBAL__N_18k_expandILi{L}ELi{mode}EEEvNS0_5GViewENS_5KPlanENS0_10ExpandArgsE"
        check_block_loop(function_body(asm, name), f"k_expand<{L}, {mode}>")
    for L in (1, 3, 4):
        name = f"_ZN4cuts12_GLOBAL__N_18k_commonILi{L}EEEvNS0_5GViewENS_5KPlanENS0_10ExpandArgsENS0_10CommonArgsE"
        check_block_loop(function_body(asm, name), f"k_common<{L}>")
    for L in (3, 4):
        name = f"_ZN4cuts12_GLOBAL__N_111k_chain_sibILi{L}EEEvNS0_5GViewENS_5KPlanENS0_7SibArgsENS0_11ChainTablesE"
        check_block_loop(function_body(asm, name), f"k_chain_sib<{L}>")


def barrier_loops(lines):
    """{header label: (depth, first branch, first exec write before it)} of
    every loop whose body (inner loops included) holds a block barrier, from
    the loop-nest annotations LLVM writes next to block labels ("Parent
    Loop", "Loop Header", "in Loop: Header")."""
    bl, cur = [], None
    for ln in lines:
        m = re.match(r"^(\.L(BB\w+)|; %bb\.\d+):(.*)$", ln)
        if m:
            cur = {"label": m.group(2) or m.group(1), "notes": [m.group(3)], "ins": []}
            bl.append(cur)
        elif cur is not None:
            st = ln.strip()
            if st.startswith(";"):
                if not cur["ins"]:
                    cur["notes"].append(st)
            elif st and not st.startswith("."):
                cur["ins"].append(st)
    for b in bl:
        chain, own = [], None
        for n in b["notes"]:
            chain += re.findall(r"Parent Loop (BB\w+) Depth=\d+", n)
            if re.search(r"This (?:Inner )?Loop Header: Depth=\d+", n):
                own = b["label"]
            h = re.search(r"in Loop: Header=(BB\w+) Depth=\d+", n)
            if h:
                own = h.group(1)
        b["chain"] = chain + ([own] if own and own not in chain else [])
    heads = {h for b in bl if any(i.startswith("s_barrier") for i in b["ins"]) for h in b["chain"]}
    out = {}
    for h in heads:
        hb = next(b for b in bl if b["label"] == h)
        br = ex = None
        for ins in hb["ins"]:
            if ins.startswith(("s_cbranch", "s_branch")):
                br = ins
                break
            if "exec" in ins and ex is None:
                ex = ins
        out[h] = (len(hb["chain"]), br, ex)
    return out


@pytest.mark.parametrize("width", [7, 11, 15])
def test_persistent_batch_loops_are_wave_uniform(tmp_path, width):
    """The persistent kernel's level loop and item loop (the two outer loops
    around its barriers, the device-wide barrier's wait included) branch on
    scalar conditions before touching the exec mask, its first barrier sits
    in the level loop, and the device-wide barrier's fences write back (release)
    and invalidate (acquire) the XCDs' L2 caches."""
    asm = assembly(os.path.join(CSRC, "batch.hip"), tmp_path)
    name = f"_ZN4cuts11k_batch_allILi{width}EEEvNS_9BatchArgsEj"
    lines = function_body(asm, name)
    loops = barrier_loops(lines)
    for depth in (1, 2):
        outer = [(h, v) for h, v in loops.items() if v[0] == depth]
        assert len(outer) == 1, f"k_batch_all<{width}>: barrier loops at depth {depth}: {outer}"
        h, (_, br, ex) = outer[0]
        assert br is not None and br.startswith(("s_cbranch_scc", "s_cbranch_vcc", "s_branch")), (h, br)
        assert ex is None, f"k_batch_all<{width}>: loop header {h} writes exec before branching ({ex})"
    first = next(b for b in blocks(lines) for ins in b[3] if ins.startswith("s_barrier"))
    assert first[1] == 1, f"k_batch_all<{width}>: first barrier at loop depth {first[1]}"
    body = [ln.strip() for ln in lines]
    assert any(ln.startswith("buffer_wbl2 sc1") for ln in body)
    assert any(ln.startswith("buffer_inv sc1") for ln in body)


@pytest.mark.parametrize("src", ["batch.hip", "match.hip"])
def test_no_sign_extended_wave_broadcasts(tmp_path, src):
    """`__builtin_amdgcn_readfirstlane` returns int; a 64-bit broadcast that
    widens its low half without a uint32 cast sign-extends it into the high
    word (seen as per-wave count sums off by multiples of 2^32).  In the
    compiler's IR that is a 32-bit shift left followed by an exact
    arithmetic shift right of an i64; no kernel may contain one."""
    out = tmp_path / (src + ".ll")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", "-emit-llvm", src, "-o", str(out)],
                   check=True, capture_output=True, cwd=CSRC)
    ir = out.read_text()
    assert re.search(r"ashr exact i64 %\w+, 32", ir) is None
