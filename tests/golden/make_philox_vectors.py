"""Generate tests/golden/philox_rocrand.json: Philox4x32-10 outputs from the
rocRAND header (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), compiled
for the host with hipcc.  rocrand_init(seed, subsequence, 0) followed by
rocrand4() evaluates Philox4x32-10 at counter {0, 0, subseq_lo, subseq_hi}
with key {seed_lo, seed_hi}; successive rocrand4() calls bump counter word 0.
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r'''
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 0xdeadbeefcafef00dull, 0x5EED0001ull, 0xffffffffffffffffull};
    const unsigned long long subs[] = {0ull, 1ull, 7ull, 0x100000000ull, 123456789012ull};
    for (unsigned long long s : seeds)
        for (unsigned long long q : subs) {
            rocrand_state_philox4x32_10 st;
            rocrand_init(s, q, 0, &st);
            for (int k = 0; k < 3; k++) {
                uint4 r = rocrand4(&st);
                std::printf("%llu %llu %d %u %u %u %u\n", s, q, k, r.x, r.y, r.z, r.w);
            }
        }
    return 0;
}
'''


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.split("\n"):
        if not line.strip():
            continue
        s, q, k, *w = line.split()
        s, q, k = int(s), int(q), int(k)
        ctr = [k, 0, q & 0xFFFFFFFF, q >> 32]
        key = [s & 0xFFFFFFFF, s >> 32]
        rows.append({"ctr": ctr, "key": key, "out": [int(x) for x in w]})
    json.dump({"source": "rocRAND rocrand_philox4x32_10.h (host build)", "vectors": rows},
              open(os.path.join(HERE, "philox_rocrand.json"), "w"), indent=0)
    print(len(rows), "vectors")


if __name__ == "__main__":
    main()
