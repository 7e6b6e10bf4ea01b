"""Runs the C++ programs in tests/cpp (Grid-API facade over libgx; they mirror Grid's own
tests/core/Test_wilson_even_odd.cc, Test_mobius_even_odd.cc, tests/solver/Test_wilson_cg_schur.cc,
Test_dwf_relupcg_prec.cc, Test_staggered_block_cg_prec.cc and Test_staggered_multishift.cc (with the
Wilson operator), Test_mobius_bcg_nosplit.cc, tests/IO/Test_nersc_io.cc, Test_ildg_io.cc and
tests/Test_dwf_mixedcg_prec.cc, plus the batched mixed-precision CG) on the GPU.  Each program asserts its own thresholds and
exits non-zero on failure."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, 'tests', 'cpp', 'bin')
CASES = [('Test_wilson_even_odd', '8.8.8.8'), ('Test_wilson_even_odd', '4.6.8.6'),
         ('Test_mobius_even_odd', '8.8.4.4'), ('Test_wilson_cg_schur', '8.8.8.8'),
         ('Test_dwf_mixedcg_prec', '8.8.8.8'), ('Test_dwf_relupcg_prec', '8.8.8.8'),
         ('Test_wilson_block_cg', '8.8.8.8'), ('Test_wilson_multishift', '8.8.8.8'),
         ('Test_nersc_io', '8.8.8.8'), ('Test_wilson_dhopdir', '8.8.4.8'), ('Test_mobius_bcg_vec', '4.4.4.8'),
         ('Test_wilson_mixedcg_batched', '8.8.8.8'), ('Test_ildg_io', '8.8.8.16')]


@pytest.mark.parametrize('name,grid', CASES)
def test_facade_program(name, grid):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f'{exe} not built (run __graft_entry__.build())'
    print(f'running {name} --grid {grid}', flush=True)
    p = subprocess.run([exe, '--grid', grid], capture_output=True, text=True, timeout=120)
    print(p.stdout[-3000:])
    print(p.stderr[-2000:])
    assert p.returncode == 0 and 'PASSED' in p.stdout
