"""The substep shape tuner's host logic (paint-sim-3d_amd/csrc/shape_tuner.h),
compiled with the host compiler into a small driver: size buckets, the
exploration order, the end of learning (also when a fused issue runs as
separate passes), the fallback of an unsampled size class to the nearest
sampled one, the periodic re-timing and a film's pinned shape.  The tuner
only picks between launch shapes that give the same bits, so these are
checks of the choice, not of any result."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "paint-sim-3d_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include "shape_tuner.h"
using czi::ShapeTuner;
using czi::FilmTuner;
namespace czi { ShapeTuner* device_substep_tuner(int) { return nullptr; } }

static void explore(ShapeTuner& t, uint64_t items, int n, bool reissue_fused) {
    std::printf("explore");
    for (int i = 0; i < n; i++) {
        int bk = 0; bool timed = false;
        int s = t.choose(items, true, bk, timed);
        // the host runs a lone fused issue (shape 5) as shape 4
        if (reissue_fused && s == 5 && timed) { t.reissue(bk, 5, 4); reissue_fused = false; }
        std::printf(" %d:%d", s, timed ? 1 : 0);
    }
    std::printf("\n");
}

int main() {
    std::printf("bucket %d %d %d %d %d %d\n", ShapeTuner::bucket(0), ShapeTuner::bucket(1), ShapeTuner::bucket(2),
                ShapeTuner::bucket(3), ShapeTuner::bucket(1024), ShapeTuner::bucket(1ull << 40));
    {   // exploration order and the end of learning
        ShapeTuner t; t.ns = 7;
        std::printf("learning0 %d\n", t.learning(1024) ? 1 : 0);
        explore(t, 1024, 22, false);
        std::printf("learning1 %d\n", t.learning(1024) ? 1 : 0);
    }
    {   // a fused issue run as shape 4 still ends the learning
        ShapeTuner t; t.ns = 7;
        explore(t, 1024, 21, true);
        std::printf("learning2 %d\n", t.learning(1024) ? 1 : 0);
        int bk = 0; bool timed = false;
        int s = t.choose(1024, true, bk, timed);
        std::printf("after %d %d\n", s, timed ? 1 : 0);
    }
    {   // an unsampled bucket takes the nearest sampled bucket's fastest shape
        ShapeTuner t; t.ns = 7;
        for (int k = 0; k < 7; k++) t.record(10, k, k == 3 ? 10.0 : 20.0, 1.0);
        for (int k = 0; k < 7; k++) t.record(5, k, k == 2 ? 10.0 : 20.0, 1.0);
        int bk = 0; bool timed = false;
        std::printf("fallback");
        for (uint64_t items : {1ull << 10, 1ull << 5, 1ull << 7, 1ull << 8, 1ull << 15, 1ull})
            std::printf(" %d", t.choose(items, false, bk, timed));
        std::printf("\n");
        ShapeTuner e; e.ns = 7;
        std::printf("empty %d\n", e.choose(1000, false, bk, timed));
    }
    {   // best sample per item; one other shape re-timed every RECHECK uses
        ShapeTuner t; t.ns = 3;
        explore(t, 64, 9, false);
        t.record(6, 0, 30.0, 2.0); t.record(6, 1, 20.0, 2.0); t.record(6, 1, 40.0, 2.0); t.record(6, 2, 25.0, 2.0);
        int bk = 0; bool timed = false, any = false;
        std::printf("recheck");
        for (int i = 1; i <= 2 * ShapeTuner::RECHECK; i++) {
            int s = t.choose(64, true, bk, timed);
            if (timed || s != 1) { std::printf(" %d:%d:%d", i, s, timed ? 1 : 0); any = true; }
        }
        std::printf("%s\n", any ? "" : " none");
    }
    {   // a film's pinned shape
        ShapeTuner t; t.ns = 7;
        FilmTuner f; f.t = &t; f.fixed = 2;
        int bk = 0; bool timed = true;
        int s = f.choose(300, true, bk, timed);
        std::printf("fixed %d %d %d %d\n", s, timed ? 1 : 0, bk, f.learning(300) ? 1 : 0);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("tuner")
    src, exe = d / "driver.cpp", d / "driver"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lines = {}
    for line in out.splitlines():
        key, _, rest = line.partition(" ")
        lines.setdefault(key, []).append(rest.split())
    return lines


def test_buckets(run):
    assert run["bucket"][0] == ["0", "0", "1", "1", "10", "23"]


def test_each_shape_issued_three_times_in_order(run):
    first = run["explore"][0]
    want = [f"{k}:1" for k in range(7) for _ in range(3)]
    assert first[:21] == want
    # the next substep takes the fallback (no sample folded yet: shape 0), untimed
    assert first[21] == "0:0"
    assert run["learning0"][0] == ["1"] and run["learning1"][0] == ["0"]


def test_fused_issue_run_as_separate_passes_still_ends_learning(run):
    assert run["learning2"][0] == ["0"]
    assert run["after"][0][1] == "0", "a further exploration was issued"


def test_unsampled_bucket_takes_nearest_sampled(run):
    # buckets 10 (shape 3 fastest) and 5 (shape 2 fastest): 10 -> 3, 5 -> 2,
    # 7 -> 2 (5 is nearer), 8 -> 3 (10 is nearer), 15 -> 3, 0 -> 2
    assert run["fallback"][0] == ["3", "2", "2", "3", "3", "2"]
    assert run["empty"][0] == ["0"]


def test_best_sample_and_periodic_recheck(run):
    # best per item: shape 1 (20/2); every RECHECK-th use re-times another
    # shape in turn and every other use takes shape 1
    rec = [tuple(int(v) for v in x.split(":")) for x in run["recheck"][0] if x != "none"]
    assert [i for i, _, _ in rec] == [64, 128]
    assert all(t == 1 and s != 1 for _, s, t in rec)
    assert rec[0][1] != rec[1][1]


def test_pinned_shape(run):
    assert run["fixed"][0] == ["2", "0", str(8), "0"]
