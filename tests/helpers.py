"""Programs the tests share: the reference's own programs from tests/golden/programs, and small
adders written with the program writer."""
import tesseract_amd as t
from tesseract_amd.programs import golden  # noqa: F401  (re-exported for the tests)


def life_cube(w, h, rbits, cbits):
    """The reference's gen_life_cube(w, h, rbits, cbits) program and its board map."""
    p = golden("life_cube_%d_%d_%d_%d" % (w, h, rbits, cbits))
    return p, p.extra["board"]


def adder(n):
    """An n-bit ripple-carry adder with carry in, shaped like the reference's examples/adder8.asm:
    inputs a<n-1>..a0 b<n-1>..b0 c, outputs s<n-1>..s0 c (the carry out replaces the carry in),
    one 19-instruction full adder per bit."""
    b = t.Builder()
    a = [b.cell("a%d" % i) for i in range(n)]
    x = [b.cell("b%d" % i) for i in range(n)]
    s = [b.cell("s%d" % i) for i in range(n)]
    c = b.cell("c")
    for cell in a[::-1] + x[::-1] + [c]:
        b.input(cell)
    for i in range(n):
        b.full_adder(s[i], c, a[i], x[i], c)
    for cell in s[::-1] + [c]:
        b.output(cell)
    return b.finish()
