"""Writes tests/golden/rng_kat.json: known-answer vectors of glibc's erand48 and
rand_r, called through ctypes on the system C library, on the aliased
uint16_t[3] state that tinypt's cpu::RandEngine uses (src/tinypt/cpu/rand.h:10-50).

Each case: a 64-bit seed mapped to the state as RandEngine(uint64_t) does, then
an op sequence (0 = (float)erand48(state), 1 = rand_r((unsigned*)state)) and the
values glibc returns.  Run: python tests/golden/make_rng_kat.py
"""
import ctypes
import json
import os

libc = ctypes.CDLL("libc.so.6")
libc.erand48.restype = ctypes.c_double
libc.erand48.argtypes = [ctypes.POINTER(ctypes.c_uint16 * 3)]
libc.rand_r.restype = ctypes.c_int
libc.rand_r.argtypes = [ctypes.c_void_p]


def state_from_seed(seed):
    s0 = seed & 0xFFFF
    s1 = (seed >> 16) & 0xFFFF
    s2 = (((seed >> 32) & 0xFFFF) + ((seed >> 48) & 0xFFFF)) & 0xFFFF
    return (ctypes.c_uint16 * 3)(s0, s1, s2)


def run(seed, ops):
    st = state_from_seed(seed)
    out = []
    for op in ops:
        if op == 0:
            out.append(float(ctypes.c_float(libc.erand48(ctypes.byref(st))).value))
        else:
            out.append(int(libc.rand_r(ctypes.addressof(st))))
    return out


def main():
    cases = []
    seeds = [0, 1, 0xDEADBEEF, (7 << 32) | 123456789, (1023 << 32) | 0xFFFFFFFF, 0xFFFFFFFFFFFFFFFF, 0x5EED5EED5EED5EED]
    for k, seed in enumerate(seeds):
        ops = [(i * 7 + k) % 3 == 0 and 1 or 0 for i in range(64)]
        cases.append({"seed": seed, "ops": ops, "values": run(seed, ops)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_kat.json")
    with open(path, "w") as f:
        json.dump({"source": "glibc erand48 / rand_r via ctypes", "cases": cases}, f, indent=0)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
