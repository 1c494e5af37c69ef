"""Known-answer vectors for Philox4x32-10 from ROCm's open-source rocRAND header.

rocRAND (MIT licence) ships its Philox4x32-10 as a header,
/opt/rocm/include/rocrand/rocrand_philox4x32_10.h.  This script compiles a
tiny host program against that header (hipcc, host code only, no GPU), which
exposes the engine's ten-round bijection and prints its output for a fixed set
of (counter, key) pairs.  The result, tests/golden/philox_kat.json, pins both
this repository's oracle and its device implementation.

Usage: python tests/golden/make_philox_kat.py
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 apply(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  static const unsigned int cases[][6] = {
    {0u, 0u, 0u, 0u, 0u, 0u},
    {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
    {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
    {1u, 0u, 0u, 0u, 0u, 0u},
    {0u, 1u, 0u, 0u, 0u, 0u},
    {0u, 0u, 0u, 1u, 0u, 0u},
    {0u, 0u, 0u, 0u, 1u, 0u},
    {0u, 0u, 0u, 0u, 0u, 1u},
    {7u, 12345u, 359u, 63u, 20261202u, 0u},
    {123456789u, 99999999u, 17u, 3u, 0xdeadbeefu, 0xcafef00du},
  };
  Probe p;
  std::printf("[");
  for (unsigned i = 0; i < sizeof(cases) / sizeof(cases[0]); ++i) {
    const unsigned* c = cases[i];
    uint4 r = p.apply(make_uint4(c[0], c[1], c[2], c[3]), make_uint2(c[4], c[5]));
    std::printf("%s{\"ctr\":[%u,%u,%u,%u],\"key\":[%u,%u],\"out\":[%u,%u,%u,%u]}", i ? "," : "",
                c[0], c[1], c[2], c[3], c[4], c[5], r.x, r.y, r.z, r.w);
  }
  std::printf("]\n");
}
"""


def main():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "kat.cpp")
        exe = os.path.join(tmp, "kat")
        open(src, "w").write(SRC)
        subprocess.run(["hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", "-I/opt/rocm/include", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vectors = json.loads(out)
    doc = {"source": "rocrand_philox4x32_10.h ten_rounds (ROCm 7.2, MIT licence)", "vectors": vectors}
    path = os.path.join(HERE, "philox_kat.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    print(path, len(vectors), "vectors")


if __name__ == "__main__":
    main()
