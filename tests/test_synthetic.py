"""CPU tests of the synthetic input generators that stand in for dray's own drivers."""
import os
import subprocess

import numpy as np
import pytest

from devil_ray_amd import synthetic as syn

_FURNACE_CPP = r"""
#include <cstdio>
#include <random>
int main(int argc, char **argv)
{
  const int n = atoi(argv[1]);
  const float b[6] = {-0.25f, 1.5f, 0.f, 1.f, 2.f, 2.125f};
  // as utilities/furnace/point_location.cpp:49-68 draws its query points
  std::linear_congruential_engine<std::uint_fast32_t, 48271, 0, 2147483647> rgen{ 0 };
  std::uniform_real_distribution<float> dx{ b[0], b[1] }, dy{ b[2], b[3] }, dz{ b[4], b[5] };
  for (int i = 0; i < n; ++i)
  {
    const float x = dx (rgen), y = dy (rgen), z = dz (rgen);
    fwrite (&x, 4, 1, stdout); fwrite (&y, 4, 1, stdout); fwrite (&z, 4, 1, stdout);
  }
  return 0;
}
"""


def test_furnace_points_match_libstdcxx(tmp_path):
    """syn.furnace_points equals what the C++ standard library's minstd_rand + uniform_real_distribution<float>
    (this machine's libstdc++, compiled with g++) draws for the furnace driver, bit for bit, over 300,000
    points (the draws past 2^16 engine outputs exercise the blocked jump-ahead)."""
    src = tmp_path / "furnace_points.cpp"
    src.write_text(_FURNACE_CPP)
    exe = tmp_path / "furnace_points"
    subprocess.check_call(["g++", "-O2", "-std=c++14", str(src), "-o", str(exe)])
    n = 300_000
    out = subprocess.run([str(exe), str(n)], check=True, capture_output=True).stdout
    ref = np.frombuffer(out, np.float32).reshape(n, 3)
    got = syn.furnace_points(n, [-0.25, 1.5, 0.0, 1.0, 2.0, 2.125])
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_minstd_known_answer():
    """The standard's required check value: the 10000th output of a default-seeded minstd_rand is 399268537
    ([rand.predef]); a seed of 0 behaves as seed 1 for this engine, like the default seed."""
    assert int(syn.minstd_outputs(10000, seed=1)[-1]) == 399268537
    assert np.array_equal(syn.minstd_outputs(1000, seed=0), syn.minstd_outputs(1000, seed=1))


def test_elements_from_nets_layout():
    """box_net gives the lattice net of an axis-aligned box, u fastest (dof i + (p+1)(j + (p+1)k))."""
    net = syn.box_net((1, 2, 3), (2, 4, 7), 2)
    assert net.shape == (27, 3)
    assert np.allclose(net[1], (1.5, 2, 3)) and np.allclose(net[3], (1, 3, 3)) and np.allclose(net[9], (1, 2, 5))
    m = syn.elements_from_nets([net, net + 1], 2)
    assert m.n_el == 2 and np.array_equal(m.elem_dofs(1), np.arange(27, 54))
