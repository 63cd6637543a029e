"""The FFT pencil tiles' LDS layout (pbh_amd/csrc/fft.hpp), checked on the host with the header itself.

The Stockham passes address a butterfly's RADIX elements as one base address plus compile-time
offsets (`lds_adv`) wherever `lds_rd_adv` / `lds_wr_adv` allow it; this compiles a host program
against fft.hpp that walks every pass of every 2^a 3^b transform size up to 1024, for every largest
radix, and checks that those offsets equal `lds_at` of the rows they stand for, and that `lds_at`
maps a tile's elements one-to-one into `lds_size` elements.  (Sizes with 9 | NC are where the
write rule has to refuse: a radix-3 pass after fewer than 8 points of powers of two.)
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SRC = r"""
#include "fft.hpp"
#include <cstdio>
#include <vector>
using namespace pm3d;

// the pass sequence of lds_fft: radix 2^ML2 passes, then the remaining power of two, then radix 3
template <typename R, int P, int NC, int ML2, int NS = 1, int REM2 = count_factor(NC, 2), int REM3 = count_factor(NC, 3)>
long passes() {
    if constexpr (REM2 == 0 && REM3 == 0) {
        return 0;
    } else {
        constexpr int L2 = REM2 >= ML2 ? ML2 : REM2;
        constexpr int RADIX = REM2 > 0 ? (1 << L2) : 3;
        constexpr int R2 = REM2 > 0 ? REM2 - L2 : 0, R3 = REM2 > 0 ? REM3 : REM3 - 1;
        constexpr int M = NC / RADIX;
        long bad = 0;
        for (int j = 0; j < M; ++j)
            for (int c = 0; c < P; ++c) {
                if (lds_rd_adv<R, P, M>())
                    for (int r = 0; r < RADIX; ++r)
                        bad += lds_at<R, P>(j + r * M, c) != lds_at<R, P>(j, c) + lds_adv<R, P>(r * M);
                const int dd = (j / NS) * NS * RADIX + j % NS;
                if (lds_wr_adv<R, P, NS, RADIX>())
                    for (int r = 0; r < RADIX; ++r)
                        bad += lds_at<R, P>(dd + r * NS, c) != lds_at<R, P>(dd, c) + lds_adv<R, P>(r * NS);
            }
        return bad + passes<R, P, NC, ML2, NS * RADIX, R2, R3>();
    }
}

template <typename R, int P, int NC>
long tile() {
    long bad = 0;
    std::vector<int> seen(lds_size<R, P>(NC), 0);
    for (int j = 0; j < NC; ++j)
        for (int c = 0; c < P; ++c) {
            const int e = lds_at<R, P>(j, c);
            if (e < 0 || e >= (int)seen.size() || seen[e]++) ++bad;
        }
    return bad + passes<R, P, NC, 1>() + passes<R, P, NC, 2>() + passes<R, P, NC, 3>() + passes<R, P, NC, 4>();
}

template <typename R, int P>
long sizes() {
    return tile<R, P, 2>() + tile<R, P, 3>() + tile<R, P, 4>() + tile<R, P, 6>() + tile<R, P, 8>() +
           tile<R, P, 12>() + tile<R, P, 16>() + tile<R, P, 24>() + tile<R, P, 32>() + tile<R, P, 48>() +
           tile<R, P, 64>() + tile<R, P, 96>() + tile<R, P, 128>() + tile<R, P, 192>() + tile<R, P, 256>() +
           tile<R, P, 384>() + tile<R, P, 512>() + tile<R, P, 768>() + tile<R, P, 1024>() + tile<R, P, 9>() +
           tile<R, P, 18>() + tile<R, P, 27>() + tile<R, P, 36>() + tile<R, P, 54>() + tile<R, P, 72>() +
           tile<R, P, 108>() + tile<R, P, 144>() + tile<R, P, 216>() + tile<R, P, 288>() + tile<R, P, 576>() +
           tile<R, P, 864>();
}

int main() {
    static_assert(lds_grouped<float, 8>() && !lds_grouped<double, 8>() && !lds_grouped<float, 16>(), "grouped tiles");
    const long bad = sizes<float, 8>() + sizes<float, 4>() + sizes<float, 16>() + sizes<double, 8>() + sizes<double, 4>();
    std::printf("%ld\n", bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(not shutil.which(HIPCC) and not Path(HIPCC).exists(), reason="hipcc not available")
def test_lds_offsets_match_layout(tmp_path):
    src = tmp_path / "lds_layout.hip"
    exe = tmp_path / "lds_layout"
    src.write_text(SRC)
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O1",
                    f"-I{ROOT / 'pbh_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "0", out.stdout + out.stderr


def test_lds_offset_rule_fails_where_it_does_not_hold():
    # the grouped layout's group term is not additive across a group end: row 7 + 1 crosses one,
    # so a constant offset from row 7 would be wrong (what lds_rd_adv / lds_wr_adv exclude)
    def at(row, c):  # rows of 8 elements in groups of 8 rows, 8 elements of padding per group
        return row * 8 + (row >> 3) * 8 + c

    def adv(d):
        return d * 8 + (d >> 3) * 8

    assert at(7 + 1, 0) != at(7, 0) + adv(1)
    assert at(3 + 8, 0) == at(3, 0) + adv(8)
