"""A second, independent operation count of the oracle's fill and pattern
sums, made by the compiler instead of by hand.

oracle/necpp_oracle.c is compiled with clang at -O0 to LLVM IR (every source
operation stays one IR instruction), and every basic block of the counted
functions gets a call that adds the block's floating-point operations to
four counters: fadd, fsub, fmul, fdiv and sqrt are flops (fdiv and sqrt also
on their own), calls of sin, cos and log are specials.  A call of one of the
file's small complex helpers (cadd, cmul, ...) adds the helper's own
operations at the call site; a call of the C99 complex division or square
root adds 9 flops (3 divisions) and 9 flops (1 division, 2 square roots), the
convention of the hand tally.  Products of constants (a literal times PI_T)
count nothing.  Comparing the two counts pins the hand-kept TALLY weights:
a weight that is off, or a branch whose operations were not tallied, shows
up as a difference.
"""
import ctypes
import os
import re
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "necpp_oracle.c")

# functions whose blocks are counted: the fill's pair loop and everything it
# calls, and the pattern's segment sum
COUNTED = {"ora_fill", "efld_compute", "eksc", "ekscx", "gx", "gxx", "gf", "intx", "conv_test", "ground_image",
           "ffld_sum"}
# helpers whose operations are added where they are called
HELPERS = {"mk", "cadd", "csub", "cneg", "cmul", "cmulr", "cdivr", "to_c", "from_c"}
SPECIAL = ("@sin(", "@cos(", "@log(")
LIBCALL = {"@__divdc3(": (9, 3, 0), "@csqrt(": (9, 1, 2)}

RT = r"""
#include <stdint.h>
static int64_t acc[4];
void ir_tally_add(int64_t f, int64_t d, int64_t q, int64_t s) { acc[0] += f; acc[1] += d; acc[2] += q; acc[3] += s; }
void ir_tally_get(int64_t* out) { for (int i = 0; i < 4; i++) out[i] = acc[i]; }
void ir_tally_reset(void) { for (int i = 0; i < 4; i++) acc[i] = 0; }
"""


def clang():
    for c in ("/opt/rocm/llvm/bin/clang", shutil.which("clang")):
        if c and os.path.exists(c):
            return c
    return None


def _functions(ir):
    """{name: [lines]} of the IR's function definitions."""
    out, cur, name = {}, None, None
    for line in ir.splitlines():
        if line.startswith("define "):
            name = re.search(r"@([A-Za-z_0-9.]+)\(", line).group(1)
            cur = [line]
        elif cur is not None:
            cur.append(line)
            if line == "}":
                out[name] = cur
                cur = None
    return out


def _blocks(body):
    """Split a function body (without its define line and closing brace)
    into basic blocks: lists of (index in body, line)."""
    blocks, cur = [], []
    for k, line in enumerate(body):
        if re.match(r"^[0-9A-Za-z_.]+:", line):
            if cur:
                blocks.append(cur)
            cur = [(k, line)]
        else:
            cur.append((k, line))
    if cur:
        blocks.append(cur)
    return blocks


def _const_regs(lines):
    """Registers loaded straight from the constant PI_T."""
    regs = set()
    for _, line in lines:
        m = re.match(r"\s*(%[0-9A-Za-z_.]+) = load double, ptr @PI_T", line)
        if m:
            regs.add(m.group(1))
    return regs


def _count(lines, helper_cost, const_regs):
    f = d = q = s = 0
    for _, line in lines:
        m = re.search(r"= (fadd|fsub|fmul|fdiv) (?:[a-z]+ )*double ([^,]+), (.+?)(?:, !.*)?$", line)
        if m:
            a, b = m.group(2).strip(), m.group(3).strip()
            lit = lambda x: re.match(r"^-?[0-9]", x) is not None or x.startswith("0x")
            if (lit(a) or a in const_regs) and (lit(b) or b in const_regs):
                continue  # a product of constants
            f += 1
            d += m.group(1) == "fdiv"
            continue
        if "call" not in line:
            continue
        if "@sqrt(" in line or "@llvm.sqrt.f64(" in line:
            f += 1
            q += 1
        elif any(t in line for t in SPECIAL):
            s += 1
        else:
            for t, (ff, dd, qq) in LIBCALL.items():
                if t in line:
                    f, d, q = f + ff, d + dd, q + qq
            m = re.search(r"@([A-Za-z_0-9]+)\(", line)
            if m and m.group(1) in helper_cost:
                hf, hd, hq, hs = helper_cost[m.group(1)]
                f, d, q, s = f + hf, d + hd, q + hq, s + hs
    return f, d, q, s


def instrument(ir):
    """The IR with a counter call at the head of every counted block."""
    funcs = _functions(ir)
    helper_cost = {}
    for h in ("mk", "to_c", "from_c", "cneg", "cadd", "csub", "cmulr", "cdivr", "cmul"):
        if h in funcs:
            body = funcs[h][1:-1]
            assert len(_blocks(body)) == 1, h  # straight-line helpers only
            helper_cost[h] = _count([(0, x) for x in body], helper_cost, set())
    out = ir
    for name in COUNTED:
        lines = funcs[name]
        body = lines[1:-1]
        new = list(body)
        for blk in reversed(_blocks(body)):
            c = _count(blk, helper_cost, _const_regs(blk))
            if c == (0, 0, 0, 0):
                continue
            call = "  call void @ir_tally_add(i64 %d, i64 %d, i64 %d, i64 %d)" % c
            k0, first = blk[0]
            at = k0 + 1 if re.match(r"^[0-9A-Za-z_.]+:", first) else k0
            while at < len(body) and re.match(r"\s*%[0-9A-Za-z_.]+ = phi ", body[at]):
                at += 1  # after the block's phi nodes
            new.insert(at, call)
        out = out.replace("\n".join(lines), "\n".join([lines[0]] + new + [lines[-1]]))
    return out + "\ndeclare void @ir_tally_add(i64, i64, i64, i64)\n"


def build(workdir):
    """Compile the instrumented oracle into workdir; returns the library path."""
    cc = clang()
    ll = os.path.join(workdir, "oracle.ll")
    subprocess.run([cc, "-O0", "-S", "-emit-llvm", "-ffp-contract=off", "-std=gnu11", "-fPIC",
                    "-DCMPLX(x,y)=__builtin_complex((double)(x),(double)(y))", SRC, "-o", ll], check=True)
    inst = os.path.join(workdir, "oracle_tally.ll")
    with open(ll) as f, open(inst, "w") as g:
        g.write(instrument(f.read()))
    rt = os.path.join(workdir, "rt.c")
    with open(rt, "w") as f:
        f.write(RT)
    so = os.path.join(workdir, "liboracle_ir.so")
    subprocess.run([cc, "-O0", "-shared", "-fPIC", inst, rt, "-o", so, "-lm"], check=True)
    return so


def read(so_path):
    L = ctypes.CDLL(so_path)
    out = (ctypes.c_int64 * 4)()
    L.ir_tally_get(out)
    return dict(zip(("flops", "divs", "sqrts", "specials"), list(out)))


def reset(so_path):
    ctypes.CDLL(so_path).ir_tally_reset()
