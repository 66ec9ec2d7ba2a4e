#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the reference's own host C.

Runs `make -C oracle ref` (which compiles PROGRAM/src/{ini,io,upd,med}.c and
random_generator.cu from the reference checkout into oracle/_ref/, one binary
per lattice configuration) and then each binary with fixed non-zero seeds.
Every dump the harness writes (see oracle/ref_harness.c) is stored in
ref_<config>.npz: blobs up to 4 KiB verbatim, larger ones as their SHA-256
digest.  Also records the SHA-256 of the reference's shipped LUT file and
the beta list it was made from.

Only this container has the reference checkout; the GPU box reads the .npz
files.  Re-run after changing oracle/ref_harness.c:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("EA_REFERENCE", "/root/reference/PROGRAM")

RAW_LIMIT = 4096


def betas_linear(tmin, tmax, nt):
    t = np.linspace(tmin, tmax, nt)
    return sorted((1.0 / t).tolist(), reverse=True)


# name: (L, Lz, NB, bc, variant, mattis, nbits, seeds(J,u,MC,Cl), jobID, betas)
CONFIGS = {
    "L8_Lz16_NB4_PBC": (8, 16, 4, 0, 0, 0, 3, (12345, 6789, 424242, 99), 5,
                        [1.4, 1.1, 0.9, 0.7]),
    "L8_Lz16_NB4_OBC": (8, 16, 4, 1, 0, 0, 2, (777, 888, 999, 1111), 0,
                        [1.5, 1.2, 0.8, 0.55]),
    "L8_Lz24_NB3_OBC_WEAK": (8, 24, 3, 1, 1, 0, 2, (31, 41, 59, 26), 1,
                             [1.2, 0.9, 0.6]),
    "L16_Lz8_NB3_OBC_STRONG": (16, 8, 3, 1, 2, 0, 2, (2718, 2818, 2845, 9045),
                               3, [1.3, 1.0, 0.75]),
    "L8_Lz16_NB4_PBC_MATTIS": (8, 16, 4, 0, 0, 1, 2, (5, 6, 7, 8), 2,
                               [1.4, 1.1, 0.9, 0.7]),
    "L8_Lz16_NB4_OBC_MATTIS": (8, 16, 4, 1, 0, 1, 2, (9, 10, 11, 12), 4,
                               [1.4, 1.1, 0.9, 0.7]),
    "L16_Lz64_NB24_OBC": (16, 64, 24, 1, 0, 0, 2, (20240601, 20240602, 20240603,
                                                   20240604), 0,
                          betas_linear(0.6, 1.6, 24)),
}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref",
                    "REF=" + REF], check=True, stdout=subprocess.DEVNULL)
    refdir = os.path.join(ROOT, "oracle", "_ref")
    for name, (L, Lz, nb, bc, var, mattis, nbits, seeds, job, betas) in \
            CONFIGS.items():
        with tempfile.TemporaryDirectory() as tmp:
            bfile = os.path.join(tmp, "betas")
            with open(bfile, "w") as f:
                for b in betas:
                    f.write("%.17g\n" % b)
            out = os.path.join(tmp, "out")
            os.mkdir(out)
            cmd = [os.path.join(refdir, "ref_" + name), out, bfile, str(nbits)]
            cmd += [str(s) for s in seeds] + [str(job)]
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
            # the betas as the reference read them (fscanf "%lf")
            arrays = {
                "meta": np.array(json.dumps({
                    "L": L, "Lz": Lz, "nt": nb, "bc": bc, "variant": var,
                    "mattis": mattis, "nbits": nbits, "nr": 8,
                    "seeds": list(seeds), "job": job}).encode()),
                "betas": np.array(betas, dtype=np.float64),
            }
            for fn in sorted(os.listdir(out)):
                if not fn.endswith(".bin"):
                    continue
                blob = open(os.path.join(out, fn), "rb").read()
                key = fn[:-4]
                if len(blob) <= RAW_LIMIT:
                    arrays["raw:" + key] = np.frombuffer(blob, dtype=np.uint8)
                else:
                    arrays["sha:" + key] = np.frombuffer(
                        hashlib.sha256(blob).digest(), dtype=np.uint8)
                    arrays["len:" + key] = np.array(len(blob), dtype=np.int64)
        path = os.path.join(HERE, "ref_%s.npz" % name)
        np.savez_compressed(path, **arrays)
        print("wrote", os.path.relpath(path, ROOT), os.path.getsize(path))

    # the shipped LUT (a data file of the reference) and its beta list
    lut = open(os.path.join(REF, "input", "LUT_for_PRNG_nbits12_NB40.bin"),
               "rb").read()
    betas40 = [float(x) for x in open(os.path.join(
        REF, "input", "betas_T07_T20_N40")).read().split()]
    with open(os.path.join(HERE, "lut_nb40.json"), "w") as f:
        json.dump({"betas": betas40,
                   "betas_text": open(os.path.join(
                       REF, "input", "betas_T07_T20_N40")).read(),
                   "sha256": hashlib.sha256(lut).hexdigest(),
                   "bytes": len(lut),
                   # a few raw thresholds for readable failures
                   "head_hex": lut[320:320 + 64].hex()}, f, indent=1)
    print("wrote tests/golden/lut_nb40.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
