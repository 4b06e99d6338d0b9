"""The device modular primitives (modarith.hpp: Shoup with the 2^63 - delta
remainder form, Moller-Granlund reduction, red2) compiled for the host with
g++ and checked against unsigned __int128 % on random and edge operands for
every prime of the 30-prime bucket."""
import os
import subprocess

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def test_modarith_primitives_exact(tmp_path):
    exe = tmp_path / "modarith_test"
    src = os.path.join(HERE, "cpp", "modarith_test.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)] + [str(p) for p in O.tables()["primes"]],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "modarith ok" in r.stdout, r.stdout + r.stderr
