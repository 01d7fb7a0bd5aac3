"""Write tests/golden/auto_tune_bloom.json: the reference's Bloom filter sizing
(auto_tune_bloom.cc, built from the reference sources by `make -C oracle ref` into
oracle/_ref/auto_tune_bloom) for a list of pq budgets.  Each entry keeps what the reference
printed: "bit <u64 words> <log2 of that> hash <k> p <false-positive rate>"."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
TOOL = os.path.join(ROOT, "oracle", "_ref", "auto_tune_bloom")
PQS = (1, 10, 32, 50, 64, 100, 128, 200, 256, 384, 512, 768, 1024, 2048, 4096)


def main():
    out = {}
    for pq in PQS:
        line = subprocess.run([TOOL, str(pq)], capture_output=True, text=True, check=True).stdout.split()
        assert line[0] == "bit" and line[3] == "hash" and line[5] == "p", line
        out[str(pq)] = {"words64": int(line[1]), "log2_words64": int(line[2]), "k": int(line[4]), "p": float(line[6])}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "auto_tune_bloom.json")
    with open(path, "w") as f:
        json.dump({"source": "reference auto_tune_bloom.cc (g++ -std=c++11)", "sizes": out}, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
