"""CopyPool (chucky_amd/csrc/host_copy.cpp), the staging copy of host-memory
appends: built with g++ beside a small driver and checked against the source
bytes for sizes and alignments around its piece, streaming and vector edges."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = r"""
#include "host_copy.hpp"
#include <cstdio>
#include <cstring>
#include <vector>
int main() {
  const size_t sizes[] = {0, 1, 15, 63, 64, 65535, 65536, 65537, (2u << 20) - 3, 2u << 20,
                          (4u << 20) + 17, (9u << 20) + 4095, 33u << 20};
  for (unsigned threads : {0u, 1u, 3u, 7u}) {
    chucky::CopyPool pool(threads);
    for (size_t n : sizes)
      for (size_t so : {0, 1, 7})
        for (size_t doff : {0, 3, 16}) {
          std::vector<unsigned char> src(n + so), dst(n + doff + 64, 0xEE);
          for (size_t i = 0; i < src.size(); ++i)
            src[i] = (unsigned char)(i * 2654435761u >> 13);
          if ((n + so + doff) % 2)
            pool.copy(dst.data() + doff, src.data() + so, n);
          else {
            pool.start(dst.data() + doff, src.data() + so, n);
            pool.wait();
          }
          if (std::memcmp(dst.data() + doff, src.data() + so, n) != 0) {
            std::printf("mismatch threads %u n %zu so %zu doff %zu\n", threads, n, so, doff);
            return 1;
          }
          for (size_t i = 0; i < doff; ++i)
            if (dst[i] != 0xEE) { std::printf("head clobbered\n"); return 1; }
          for (size_t i = n + doff; i < dst.size(); ++i)
            if (dst[i] != 0xEE) { std::printf("tail clobbered n %zu\n", n); return 1; }
        }
  }
  std::printf("ok\n");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_copy_pool_copies_every_byte(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    csrc = os.path.join(ROOT, "chucky_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", csrc, str(src),
                    os.path.join(csrc, "host_copy.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
