"""The LDS table's bucket hash (csrc/jc_hash.hpp) stays inside the table for
every key and table size: compiled for the host with g++ and checked over
edge-case and random keys. (A signed shift here once produced indices far
outside the table.)"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sycl-jaccard_amd", "csrc")

PROG = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include "jc_hash.hpp"
int main() {
  std::mt19937_64 rng(1);
  const int32_t edge[] = {0, 1, 2, 0x7FFFFF, 0x800000, 0xFFFFFF, 0x1000000, 0x7FFFFFFF, 0x40000000, 12345678};
  long bad = 0, n = 0;
  for (int lb = 0; lb <= 14; ++lb) {
    for (int32_t y : edge) { n++; if (jc::hash_bucket(y, lb) >= (1u << lb)) bad++; }
    for (int i = 0; i < 200000; ++i) {
      const int32_t y = (int32_t)(rng() & 0x7FFFFFFF);
      n++;
      if (jc::hash_bucket(y, lb) >= (1u << lb)) bad++;
    }
    // consecutive ids: every bucket index in range, and no bucket takes more
    // than a small share of a dense id block
    std::vector<int> cnt(1u << lb, 0);
    const int keys = 3 << lb;
    for (int32_t y = 1000; y < 1000 + keys; ++y) { n++; uint32_t b = jc::hash_bucket(y, lb); if (b >= (1u << lb)) bad++; else cnt[b]++; }
    int mx = 0; for (int c : cnt) mx = c > mx ? c : mx;
    printf("lb %d max bucket load of %d consecutive ids: %d\n", lb, keys, mx);
    if (lb >= 2 && mx > 6) bad++;  // 3 keys per bucket on average
  }
  printf("checked %ld bad %ld\n", n, bad);
  return bad != 0;
}
"""


def test_bucket_in_range(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text(PROG.replace("#include <random>", "#include <random>\n#include <vector>"))
    exe = tmp_path / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "bad 0" in r.stdout
