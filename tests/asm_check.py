"""Checks on the gfx950 assembly that hipcc generates for the pipelined BF16
GEMM kernel (gemm_bf16_pipe_kernel).

The kernel issues LDS fragment reads and LDS-DMA loads from inline asm and
lets the reads complete after the issuing statement ends; its accumulators are
written by MFMAs the compiler does not see.  That is sound only if the
compiler-generated code around the asm statements keeps three rules, which
these functions verify on the text of the generated assembly:

  * no instruction touches the destination registers of an LDS read until a
    s_waitcnt lgkmcnt(N) has retired that read (LDS operations complete in
    order, so lgkmcnt(N) retires all but the youngest N LDS reads);
  * no non-MFMA instruction reads or writes an accumulator register within 18
    wait states of the MFMA that wrote it (the gfx950 MFMA 16x16x32 result
    latency is 8 passes plus a few states), and no MFMA reads an accumulator
    as srcC within 3 states of a VALU write to it;
  * M0, which the asm statements overwrite, is not used by compiler-generated
    code.

The scan is linear over the function body, which is how the kernel's code is
laid out (the step loop is straight-line code between the barrier and the
loop branch).
"""
import re

REG = re.compile(r'(?<![\w.])([va])(?:\[(\d+):(\d+)\]|(\d+)(?!\w))')


def regs(ops):
    out = set()
    for m in REG.finditer(ops):
        k = m.group(1)
        if m.group(4) is not None:
            out.add((k, int(m.group(4))))
        else:
            for r in range(int(m.group(2)), int(m.group(3)) + 1):
                out.add((k, r))
    return out


def functions(text, pattern=r'_ZN4mpsa21gemm_bf16_pipe'):
    """(name, lines) of every function whose symbol starts with pattern."""
    for m in re.finditer(r'^(%s\S*):' % pattern, text, re.M):
        i = m.end()
        j = text.index('.Lfunc_end', i)
        yield m.group(1), text[i:j].split('\n')


def check(lines):
    """List of rule violations in one function body."""
    errs = []
    queue = []   # LDS reads in flight, oldest first: (line, destination registers)
    pend = {}    # accumulator -> wait states since the MFMA that wrote it
    vw = {}      # accumulator -> wait states since a VALU write
    in_asm = False
    for ln, raw in enumerate(lines):
        t = raw.strip()
        if t.startswith(';;#ASMSTART'):
            in_asm = True
            continue
        if t.startswith(';;#ASMEND'):
            in_asm = False
            continue
        line = raw.split(';')[0].strip()
        if not line or line.startswith('.') or line.endswith(':'):
            continue
        mn, _, ops = line.partition(' ')
        if mn == 's_waitcnt':
            m = re.search(r'lgkmcnt\((\d+)\)', ops)
            if m:
                n = int(m.group(1))
                while len(queue) > n:
                    queue.pop(0)
            continue
        if not in_asm and re.search(r'\bm0\b', ops):
            errs.append('line %d: compiler-generated code uses M0: %s' % (ln, line))
        rs = regs(ops)
        for (q, d) in queue:
            if rs & d:
                errs.append('line %d uses registers of the LDS read at line %d still in flight: %s' % (ln, q, line))
        states = int(ops, 0) + 1 if mn == 's_nop' else 1
        if mn.startswith('v_mfma'):
            parts = [p.strip() for p in ops.split(',')]
            for r in regs(parts[3]):
                if r in vw and vw[r] < 3:
                    errs.append('line %d: MFMA reads %s%d %d states after a VALU write' % (ln, r[0], r[1], vw[r]))
            for r in regs(parts[0]):
                pend[r] = 0
        else:
            for r in rs:
                if r[0] == 'a' and r in pend and pend[r] < 18:
                    errs.append('line %d touches %s%d %d states after an MFMA wrote it: %s'
                                % (ln, r[0], r[1], pend[r], line))
            if mn.startswith('v_accvgpr_write') or mn.startswith('v_accvgpr_mov'):
                for r in regs(ops.split(',')[0]):
                    vw[r] = 0
        for d in (pend, vw):
            for r in list(d):
                d[r] += states
                if d[r] > 64:
                    del d[r]
        if mn.startswith('ds_read'):
            queue.append((ln, regs(ops.split(',')[0])))
    return errs
