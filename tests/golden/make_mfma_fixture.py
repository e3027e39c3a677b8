"""Writes mfma_f16_32x32x16_gfx950.npz: operands and results of
v_mfma_f32_32x32x16_f16 measured on an MI355X (gfx950), the hardware side of
the learned-layer predictor's product model (oracle lb_mfma16, checked in
tests/test_learned_oracle.py).  Run on a machine with the GPU:

    python tests/golden/make_mfma_fixture.py

It builds mfma_probe.hip with hipcc, runs it (1536 cases, six operand
regimes in turn) and keeps the first case of each regime."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> int:
    with tempfile.TemporaryDirectory() as d:
        exe, out = os.path.join(d, "probe"), os.path.join(d, "mfma.bin")
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-o", exe, os.path.join(HERE, "mfma_probe.hip")], check=True)
        subprocess.run([exe, out], check=True)
        with open(out, "rb") as f:
            n = int(np.frombuffer(f.read(4), np.int32)[0])
            a = np.frombuffer(f.read(n * 512 * 2), np.float16).reshape(n, 32, 16)
            b = np.frombuffer(f.read(n * 512 * 2), np.float16).reshape(n, 16, 32)
            c = np.frombuffer(f.read(n * 1024 * 4), np.float32).reshape(n, 32, 32)
            dd = np.frombuffer(f.read(n * 1024 * 4), np.float32).reshape(n, 32, 32)
    k = 6
    np.savez_compressed(os.path.join(HERE, "mfma_f16_32x32x16_gfx950.npz"), a=a[:k].copy(),
                        b=b[:k].copy(), c=c[:k].copy(), d=dd[:k].copy())
    return 0


if __name__ == "__main__":
    sys.exit(main())
