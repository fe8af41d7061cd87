"""Check the generated gfx950 code of the library's kernels for one hazard class.

A VGPR written by a memory load must never reach a use as the upper half of a
64-bit register pair v[a:a+1] whose lower half the same load did not write. No
kernel of this library builds a 64-bit value that way (loaded 64-bit data, e.g.
complex64, is loaded as a whole pair; 64-bit addresses and indices are built by
arithmetic), so such a path means the register allocator gave the load's result
the register of a live value -- the upper half of a zero-extended index. That is
exactly the miscompile that made the F-engine address its twiddle and output
buffers 2^32 * mask_byte elements out of bounds.

The check follows every control-flow path from each load (basic blocks split at
labels and branches) until the loaded register is redefined.

`waterfall_loops` finds a second, performance-only pattern: a block that loops
over the distinct values of a scalar operand across the wave's lanes, the form
the compiler gives a buffer access whose descriptor it cannot prove uniform
(tests/test_cpu_emu.py requires none in the F-engine kernels).

usage: python tests/isa_lint.py file.s [file.s ...]   (exit 1 on a hazard finding)
"""

import re
import sys

LOAD = re.compile(r"^(global_load|flat_load|buffer_load|scratch_load|ds_read|ds_load)")
VMEM_ADDR = re.compile(r"^(global_|flat_|buffer_|scratch_)")
REG = re.compile(r"v\[(\d+):(\d+)\]|v(\d+)\b")


def regs(tok):
    """VGPR numbers named by one operand token."""
    out = []
    for m in REG.finditer(tok):
        if m.group(3) is not None:
            out.append(int(m.group(3)))
        else:
            out.extend(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def split_ops(s):
    return [t.strip() for t in s.split(",")] if s else []


def parse_kernels(text):
    """{kernel: [(label or None, mnemonic, operands)]} for every function body."""
    kernels, cur, name = {}, None, None
    for line in text.splitlines():
        line = line.split(";")[0].rstrip()
        if not line.strip():
            continue
        m = re.match(r"^([.\w$]+):", line)
        if m:
            lab = m.group(1)
            if lab.startswith(".LBB"):
                if cur is not None:
                    cur.append((lab, None, []))
            elif lab.startswith(".Lfunc_end"):
                cur = None
            elif not lab.startswith("."):
                name = lab
                cur = kernels.setdefault(name, [])
            continue
        if cur is None:
            continue
        s = line.strip()
        if s.startswith("."):
            if s.startswith(".Lfunc_end") or s.startswith(".size"):
                cur = None
            continue
        parts = s.split(None, 1)
        cur.append((None, parts[0], split_ops(parts[1] if len(parts) > 1 else "")))
        if parts[0] == "s_endpgm":
            pass
    return kernels


def blocks(insts):
    """Split into basic blocks: list of (label, [insts]) plus successor lists."""
    bl = [[None, []]]
    for lab, mn, ops in insts:
        if lab is not None:
            if bl[-1][1] or bl[-1][0] is not None:
                bl.append([lab, []])
            else:
                bl[-1][0] = lab
            continue
        bl[-1][1].append((mn, ops))
        if mn.startswith("s_branch") or mn.startswith("s_cbranch") or mn == "s_endpgm" \
                or mn.startswith("s_setpc"):
            bl.append([None, []])
    bl = [b for b in bl if b[0] is not None or b[1]]
    index = {b[0]: i for i, b in enumerate(bl) if b[0] is not None}
    succ = []
    for i, (lab, ins) in enumerate(bl):
        s = []
        last = ins[-1] if ins else None
        if last and last[0].startswith("s_branch"):
            s.append(index[last[1][0]])
        elif last and last[0].startswith("s_cbranch"):
            s.append(index[last[1][0]])
            if i + 1 < len(bl):
                s.append(i + 1)
        elif last and (last[0] == "s_endpgm" or last[0].startswith("s_setpc")):
            pass
        elif i + 1 < len(bl):
            s.append(i + 1)
        succ.append(s)
    return bl, succ


def vgpr_defs(mn, ops):
    if not ops:
        return []
    if mn.startswith(("v_cmp", "v_readfirstlane", "v_readlane")):
        return []
    if mn.startswith("v_") or LOAD.match(mn):
        return regs(ops[0])
    return []


def has_dest(mn):
    if LOAD.match(mn):
        return True
    return mn.startswith("v_")


def pair_uses(mn, ops):
    """(lower, upper) of every 64-bit VGPR pair read by the instruction."""
    srcs = ops[1:] if has_dest(mn) else ops
    out = []
    for t in srcs:
        m = re.fullmatch(r"-?v\[(\d+):(\d+)\]", t)
        if m and int(m.group(2)) == int(m.group(1)) + 1:
            out.append((int(m.group(1)), int(m.group(2))))
    return out


def lint_kernel(insts):
    bl, succ = blocks(insts)
    findings = []
    for bi, (_, ins) in enumerate(bl):
        for ii, (mn, ops) in enumerate(ins):
            if not LOAD.match(mn):
                continue
            dest = set(regs(ops[0]))
            for r in dest:
                seen = set()
                stack = [(bi, ii + 1)]
                while stack:
                    b, start = stack.pop()
                    if (b, start) in seen:
                        continue
                    seen.add((b, start))
                    killed = False
                    for k in range(start, len(bl[b][1])):
                        m2, o2 = bl[b][1][k]
                        for lo, hi in pair_uses(m2, o2):
                            if hi == r and lo not in dest:
                                findings.append((mn, ops, m2, o2))
                        if r in vgpr_defs(m2, o2):
                            killed = True
                            break
                    if not killed:
                        for s in succ[b]:
                            stack.append((s, 0))
    return findings


def waterfall_loops(insts):
    """Labels of the blocks that loop over the distinct values of a scalar
    operand across the wave's lanes (the compiler's form of a memory access
    whose descriptor or base it could not prove uniform): the block reads a
    lane's value with v_readfirstlane, masks exec to the lanes that share it
    (s_and_saveexec), and branches back to itself while lanes remain."""
    bl, _ = blocks(insts)
    out = []
    for lab, ins in bl:
        if lab is None or not ins:
            continue
        last = ins[-1]
        mns = {mn for mn, _ in ins}
        if last[0] == "s_cbranch_execnz" and last[1] and last[1][0] == lab and \
                "v_readfirstlane_b32" in mns and "s_and_saveexec_b64" in mns:
            out.append(lab)
    return out


def lint_text(text):
    out = []
    for name, insts in parse_kernels(text).items():
        for f in lint_kernel(insts):
            out.append((name, f))
    return out


def main(paths):
    bad = 0
    for p in paths:
        for name, (mn, ops, m2, o2) in lint_text(open(p).read()):
            bad += 1
            print(f"{p}: {name}: {mn} {', '.join(ops)}  reaches  {m2} {', '.join(o2)}")
    print(f"{bad} finding(s)")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
