"""Host-side checks of the closed-form offsets of the split-BF16 operand
layouts and of the FFT pass padding rule (compiled with hipcc as host code,
no GPU).

Each closed form is compared with the sum or the enumeration it stands for:
the row lengths summed element by element, every segment's start against
the running sum of the segment sizes, and every padded LDS index that an
FFT pass forms from a per-butterfly base plus a constant against the padded
index computed directly.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "exoplasim_amd", "csrc")

PROGRAM = r"""
#include <cstdio>
#include <utility>
#include "sht_common.h"
using namespace exo;

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s line %d\n", #c, __LINE__); ++fails; } } while (0)

// the padded index of in-row element e (rows of N + N/32 complex)
static int padded(int e) { return e + (e >> 5); }

template <int N, int R, int NS, int r>
static void check_radix_index() {
  constexpr int NB = N / R;
  // reads: element j + r NB from the base padded(j)
  {
    constexpr int K = r * NB;
    constexpr int D = pad_delta<NB, NS, R, K, false>();
    if (D >= 0)
      for (int j = 0; j < NB; ++j) CHECK(padded(j) + K + D == padded(j + K));
  }
  // writes: Stockham output index idxd + r NS from the base padded(idxd)
  {
    constexpr int K = r * NS;
    constexpr int D = pad_delta<NB, NS, R, K, true>();
    if (D >= 0)
      for (int j = 0; j < NB; ++j) {
        const int idxd = (j / NS) * NS * R + j % NS;
        CHECK(padded(idxd) + K + D == padded(idxd + K));
      }
  }
}
template <int N, int R, int NS, int... rs>
static void check_pass_rs(std::integer_sequence<int, rs...>) {
  (check_radix_index<N, R, NS, rs>(), ...);
}
template <int N, int R, int NS>
static void check_pass() {
  check_pass_rs<N, R, NS>(std::make_integer_sequence<int, R>{});
}

static void check_layout(int T) {
  // row lengths: 8 ceil((T+1-m)/16), and their closed-form sum
  int sum = 0;
  for (int L = 1; L <= T + 1; ++L) sum += 8 * ((L + 15) / 16);
  CHECK(pk_c16(T + 1) == sum);
  for (int x = 0; x <= T + 1; ++x) {
    int s = 0;
    for (int L = 1; L <= x; ++L) s += 8 * ((L + 15) / 16);
    CHECK(pk_c16(x) == s);
  }
  // segments: spectral [re/im 2][parity 2][rw] fp32, table [plane 3][row 16][rw]
  int pk = 0, tab = 0;
  for (int m = 0; m <= T; ++m) {
    const int rw = pk_rw(T, m);
    const int L = T + 1 - m;
    CHECK(rw == 8 * ((L + 15) / 16));
    CHECK(rw % 8 == 0 && 2 * rw >= L);   // a parity row holds ceil(L/2) values
    CHECK(pk_seg(T, m) == pk);
    CHECK(tab_seg(T, m) == tab);
    pk += 4 * rw;
    tab += 48 * rw;
    // every n - m < L lies in K block kb = (n - m) / 32, at j = (n - m) / 2
    // within [16 kb, 16 kb + 16) and below rw
    for (int d = 0; d < L; ++d) {
      const int kb = d / 32, j = d / 2;
      CHECK(j >= 16 * kb && j < 16 * kb + 16 && j < rw);
    }
  }
  CHECK(pk == pk_field(T));
  CHECK(tab == tab_pb(T));
}

int main() {
  for (int T : {21, 42, 63, 127}) check_layout(T);
  check_pass<64, 8, 1>();
  check_pass<64, 8, 8>();
  check_pass<128, 16, 1>();
  check_pass<128, 8, 16>();
  check_pass<192, 8, 1>();
  check_pass<192, 8, 8>();
  check_pass<192, 3, 64>();
  check_pass<384, 16, 1>();
  check_pass<384, 8, 16>();
  check_pass<384, 3, 128>();
  std::printf(fails ? "failed\n" : "ok\n");
  return fails ? 1 : 0;
}
"""


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    return None


def test_operand_layouts_and_fft_padding(tmp_path):
    cc = hipcc()
    if cc is None:
        pytest.skip("hipcc not available")
    src = tmp_path / "layouts.cpp"
    exe = tmp_path / "layouts"
    src.write_text(PROGRAM)
    build = subprocess.run([cc, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1",
                            "-I", CSRC, "-o", str(exe), str(src)],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), run.stdout[-3000:]
