#!/bin/bash
# Writes tests/golden/philox_rocrand_kat.json: Philox4x32-10 outputs computed
# by rocRAND's host-callable engine (rocrand_device::philox4x32_10_engine in
# /opt/rocm/include/rocrand/rocrand_philox4x32_10.h) for a set of counters and
# keys.  Counter (c0,c1,c2,c3) is reached with subsequence = c3<<32|c2 and
# offset = 4*(c1<<32|c0), so c1 < 2^30 keeps the offset in range.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
tmp="$(mktemp -d)"
cat > "$tmp/kat.cpp" <<'CPP'
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
  const unsigned cases[][6] = {
    {0, 0, 0, 0, 0, 0}, {1, 0, 0, 0, 1234, 0}, {7, 0, 3, 0x105, 1234, 0},
    {0xFFFFFFFEu, 5, 17, 0x203, 1234, 0}, {12, 99, 400, 0x7FFF06, 0xDEADBEEF, 0x12345678},
    {3, 1, 49999, 0x00000101, 1234, 0}, {0x1000, 0x3FFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF}};
  printf("[\n");
  const int n = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < n; ++i) {
    const unsigned *c = cases[i];
    unsigned long long seed = ((unsigned long long)c[5] << 32) | c[4];
    unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
    unsigned long long off = (((unsigned long long)c[1] << 32) | c[0]) * 4ull;
    rocrand_device::philox4x32_10_engine e(seed, sub, off);
    unsigned r[4];
    for (int k = 0; k < 4; ++k) r[k] = e();
    printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
           c[0], c[1], c[2], c[3], c[4], c[5], r[0], r[1], r[2], r[3], i + 1 < n ? "," : "");
  }
  printf("]\n");
}
CPP
hipcc -O1 -o "$tmp/kat" "$tmp/kat.cpp" 2>/dev/null
"$tmp/kat" > "$here/philox_rocrand_kat.json"
rm -rf "$tmp"
echo "wrote $here/philox_rocrand_kat.json"
