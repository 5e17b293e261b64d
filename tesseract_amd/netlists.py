"""Seeded random NAND netlists, NAND-level rewrites of them and injected gate faults: the equiv
circuits of the serving configuration, written as ISCAS .bench text for the netlist loader (the
reference's net.c reads the same format, src/net.c:110-357); and its verify requests, multipliers
against their spec.

circuit(n_in, n_out, gates, rng): gate k = NAND of two earlier signals with the locality logic has:
the first from the most recent 64 signals, the second a primary input (1 in 10) or one of the most
recent 256 signals; the outputs are the last n_out gates. (Operands drawn uniformly from every
earlier signal keep thousands of values live at once; such circuits compile to code that spills
and takes minutes of hipRTC time at 5,000 gates.)

rewrite(c, rng): every gate g = NAND(a, b) becomes one of these equivalent NAND structures:
  NAND(b, a)                                  (operands swapped)
  t = NAND(a, b); u = NAND(t, t); NAND(u, u)  (double inversion)
  NAND(NAND(a, NAND(b, b)), a)                (!(a & !(a & !b)) = !(a & b))
fault(c, rng): one gate of the rewrite reads a different, earlier signal in place of one input.
"""
import numpy as np


class Circuit:
    def __init__(self, inputs, gates, outputs):
        self.inputs = inputs      # names
        self.gates = gates        # (name, a, b) = NAND(a, b), topological
        self.outputs = outputs    # names

    def bench(self):
        lines = ["INPUT(%s)" % i for i in self.inputs]
        lines += ["OUTPUT(%s)" % o for o in self.outputs]
        lines += ["%s = NAND(%s, %s)" % g for g in self.gates]
        return "\n".join(lines) + "\n"

    def evaluate(self, x):
        """Outputs for assignment x (input k = bit n-1-k of x), for checks on a few points."""
        v = {}
        n = len(self.inputs)
        for k, name in enumerate(self.inputs):
            v[name] = (x >> (n - 1 - k)) & 1
        for name, a, b in self.gates:
            v[name] = 1 - (v[a] & v[b])
        return [v[o] for o in self.outputs]


def circuit(n_in, n_out, gates, rng):
    inputs = ["x%d" % i for i in range(n_in)]
    signals = list(inputs)
    out = []
    for k in range(gates):
        recent = signals[-64:]
        a = recent[int(rng.integers(0, len(recent)))]
        if rng.random() < 0.1:
            b = inputs[int(rng.integers(0, n_in))]
        else:
            wide = signals[-256:]
            b = wide[int(rng.integers(0, len(wide)))]
        name = "g%d" % k
        out.append((name, a, b))
        signals.append(name)
    return Circuit(inputs, out, [g[0] for g in out[-n_out:]])


def rewrite(c, rng):
    gates = []
    for name, a, b in c.gates:
        r = int(rng.integers(0, 3))
        if r == 0:
            gates.append((name, b, a))
        elif r == 1:
            gates += [(name + "_t", a, b), (name + "_u", name + "_t", name + "_t"),
                      (name, name + "_u", name + "_u")]
        else:
            gates += [(name + "_nb", b, b), (name + "_m", a, name + "_nb"), (name, name + "_m", a)]
    return Circuit(list(c.inputs), gates, list(c.outputs))


def fault(c, rng):
    gates = list(c.gates)
    k = int(rng.integers(len(gates) // 2, len(gates)))
    name, a, b = gates[k]
    earlier = list(c.inputs) + [g[0] for g in gates[:k]]
    other = earlier[int(rng.integers(0, len(earlier)))]
    gates[k] = (name, a, other) if other != b else (name, other, b)
    return Circuit(list(c.inputs), gates, list(c.outputs))


def request(seed, k):
    """Request k of the serving stream: (circuit, rewrite-with-maybe-fault, n_inputs, faulty)."""
    rng = np.random.default_rng([seed, k])
    n_in = int(round(np.exp(rng.uniform(np.log(8), np.log(36)))))
    n_out = int(rng.integers(1, 33))
    gates = int(round(np.exp(rng.uniform(np.log(50), np.log(20000)))))
    gates = max(gates, n_out)
    c = circuit(n_in, n_out, gates, rng)
    r = rewrite(c, rng)
    faulty = rng.random() < 0.1
    if faulty:
        r = fault(r, rng)
    return c, r, n_in, faulty


# Verify requests of the stream: arithmetic circuits (the reference's gen_mul multipliers, 4 to 10
# bits a side: input spaces 2^8..2^20) checked against their spec; one in ten carries a planted
# spec difference (a product off by one where a and b are both odd).
VERIFY_EVERY = 5


def verify_request(seed, k):
    """(program name, spec, n_inputs, planted) of verify request k."""
    rng = np.random.default_rng([seed, k, 1])
    bits = int(rng.integers(4, 11))
    planted = rng.random() < 0.1
    spec = "p = a * b + (a & b & 1)" if planted else "p = a * b"
    return "mul%d" % bits, spec, 2 * bits, planted
