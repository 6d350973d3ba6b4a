"""Writes the cache golden fixtures from the reference cache generator.

The generator is the reference's own backend/generate_bit_transition_cache.cpp,
compiled from source by `make -C oracle ref` into oracle/_ref/ (never shipped).
It is run in a scratch directory; its output files cache_{n',k'} for
0 <= k' < n' <= N_MAX are reduced to

  cache_sha256.json   {"n_k": sha256 hex} for every file (n' <= 14)
  raw/cache_n_k       the raw files themselves for n' <= 6 (about 11 KB)

Usage: python tests/golden/make_golden.py [N_MAX=14]
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(REPO, "oracle", "_ref", "generate_bit_transition_cache")
RAW_MAX = 6


def main():
    n_max = int(sys.argv[1]) if len(sys.argv) > 1 else 14
    if not os.path.exists(GEN):
        subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
    work = tempfile.mkdtemp(prefix="bdc_golden_")
    try:
        os.makedirs(os.path.join(work, "transition_counts"))
        subprocess.check_call([GEN, str(n_max)], cwd=work)
        digests = {}
        raw_dir = os.path.join(HERE, "raw")
        os.makedirs(raw_dir, exist_ok=True)
        for n in range(0, n_max + 1):
            for k in range(0, n):
                name = f"cache_{n}_{k}"
                path = os.path.join(work, "transition_counts", name)
                with open(path, "rb") as f:
                    data = f.read()
                assert len(data) == 4 << (n + k), (name, len(data))
                digests[f"{n}_{k}"] = hashlib.sha256(data).hexdigest()
                if n <= RAW_MAX:
                    shutil.copyfile(path, os.path.join(raw_dir, name))
        with open(os.path.join(HERE, "cache_sha256.json"), "w") as f:
            json.dump({"generator": "backend/generate_bit_transition_cache.cpp (reference, built from source)",
                       "n_max": n_max, "sha256": digests}, f, indent=1, sort_keys=True)
        print(f"wrote {len(digests)} digests")
    finally:
        shutil.rmtree(work)


if __name__ == "__main__":
    main()
