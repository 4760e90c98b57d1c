"""Write the golden vectors under tests/golden/ from the compiled reference.

Run in a checkout where ``oracle/_ref/libadrt_ref.so`` was built from the
reference sources (``make -C oracle``):

    python tests/golden/make_golden.py

Outputs (committed):
  golden_small.npz   seeded inputs and the reference's outputs for n <= 16
                     (adrt, adrt_step, bdrt, bdrt_step, iadrt, fmg_*,
                     interp_to_cart), f32/f64
  golden_hashes.json SHA-256 of the reference's outputs on larger seeded
                     inputs (values canonicalised: -0.0 -> +0.0)
The reference itself never enters the repository; only these vectors do.
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402


def canonical_sha256(a):
    a = np.ascontiguousarray(a)
    a = np.where(a == 0, np.zeros((), a.dtype), a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def small_vectors():
    out = {}
    for dt in ("float32", "float64"):
        for n in (1, 2, 4, 8, 16):
            rng = np.random.default_rng([7, n, 0 if dt == "float32" else 1])
            img = rng.standard_normal((2, n, n)).astype(dt)
            dat = rng.standard_normal((2, 4, 2 * n - 1, n)).astype(dt)
            key = f"{dt}_n{n}"
            out[f"{key}_img"] = img
            out[f"{key}_dat"] = dat
            out[f"{key}_adrt"] = ref.adrt(img)
            out[f"{key}_bdrt"] = ref.bdrt(dat)
            out[f"{key}_iadrt"] = ref.iadrt(dat)
            out[f"{key}_iadrt_of_adrt"] = ref.iadrt(ref.adrt(np.round(img * 8)))
            for s in range(max(n.bit_length() - 1, 0)):
                out[f"{key}_adrt_step{s}"] = ref.adrt_step(dat, s)
                out[f"{key}_bdrt_step{s}"] = ref.bdrt_step(dat, s)
            if n >= 2:
                out[f"{key}_restriction"] = ref.fmg_restriction(dat)
            hp = rng.standard_normal((2, n + 1, n + 2)).astype(dt)
            out[f"{key}_hp_in"] = hp
            out[f"{key}_highpass"] = ref.fmg_highpass(hp)
            out[f"{key}_prolongation"] = ref.fmg_prolongation(hp)
            if n >= 2:
                out[f"{key}_interp"] = ref.interp_to_cart(dat)
    return out


def hash_vectors():
    res = {}
    # Forward parity case: 8 x 128^2 float32, uniform [0, 1) from seed 0.
    x = np.random.default_rng(0).random((8, 128, 128), dtype=np.float32)
    res["adrt_f32_8x128_seed0"] = canonical_sha256(ref.adrt(x))
    x64 = np.random.default_rng(0).random((2, 256, 256))
    res["adrt_f64_2x256_seed0"] = canonical_sha256(ref.adrt(x64))
    y = ref.adrt(x)
    res["bdrt_f32_of_adrt_8x128_seed0"] = canonical_sha256(ref.bdrt(y))
    res["iadrt_f32_of_adrt_8x128_seed0"] = canonical_sha256(ref.iadrt(y))
    big = np.random.default_rng(1).random((1, 1024, 1024), dtype=np.float32)
    res["adrt_f32_1x1024_seed1"] = canonical_sha256(ref.adrt(big))
    res["interp_f32_of_adrt_8x128_seed0"] = canonical_sha256(ref.interp_to_cart(y))
    res["interp_f64_of_adrt_2x256_seed0"] = canonical_sha256(ref.interp_to_cart(ref.adrt(x64)))
    return res


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/libadrt_ref.so is not built; run `make -C oracle` first")
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small_vectors())
    with open(os.path.join(HERE, "golden_hashes.json"), "w") as f:
        json.dump(hash_vectors(), f, indent=1, sort_keys=True)
    print("wrote golden_small.npz and golden_hashes.json")


if __name__ == "__main__":
    main()
