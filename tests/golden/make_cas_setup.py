"""Writes tests/golden/cas_setup.json: CasSetup outputs of the reference's own ffx_cas.h/ffx_a.h,
compiled for the CPU into oracle/_ref/libffx_ref.so by oracle/Makefile (needs the reference tree
in this container; the JSON is what the tests read everywhere else)."""
import ctypes as C
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libffx_ref.so"))
    L.ref_cas_setup.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)] + [C.c_float] * 5
    cases = []
    for i in range(41):
        s = -0.25 + i * 0.0375
        for (w, h) in [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]:
            c0, c1 = (C.c_uint32 * 4)(), (C.c_uint32 * 4)()
            L.ref_cas_setup(c0, c1, s, w, h, w, h)
            cases.append(dict(sharpness=s, in_w=w, in_h=h, const0=list(c0), const1=list(c1)))
    out = dict(source="shaders/ffx_cas.h:375-394 CasSetup (A_CPU) built from the reference tree",
               cases=cases)
    with open(os.path.join(HERE, "cas_setup.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
