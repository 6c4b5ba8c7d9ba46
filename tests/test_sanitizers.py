"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only; GPU code is never built with them):
the oracle restatement (oracle/griffon_oracle.cpp, Makefile target asan) driven over every exported function on the
reference fixtures, and the CPU build of the device eigenvalue routine (spitfire_amd/csrc/gf_eig.h) over random,
badly scaled, triangular, defective and companion matrices. Any sanitizer report fails the run."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _env():
    asan = subprocess.run(['gcc', '-print-file-name=libasan.so'], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip('libasan is not installed')
    env = dict(os.environ)
    env['ASAN_OPTIONS'] = 'detect_leaks=0:abort_on_error=1'
    env['UBSAN_OPTIONS'] = 'halt_on_error=1:print_stacktrace=1'
    return env, asan


def _clean(r):
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert 'AddressSanitizer' not in out and 'runtime error' not in out, out[-4000:]
    return out


def test_oracle_under_address_and_undefined_behavior_sanitizers():
    env, asan = _env()
    subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'oracle'), 'asan'], check=True)
    env['GF_ORACLE_SO'] = os.path.join(ROOT, 'oracle', '_asan', 'liboracle.so')
    env['LD_PRELOAD'] = asan
    env['OMP_NUM_THREADS'] = '1'
    r = subprocess.run([sys.executable, os.path.join(HERE, 'native', 'oracle_sanitize_run.py')], env=env,
                       capture_output=True, text=True, timeout=600)
    out = _clean(r)
    assert 'chem_c1.npz ok' in out and 'chem_53.npz ok' in out


def test_eigenvalue_routine_under_sanitizers(tmp_path):
    env, _ = _env()
    exe = str(tmp_path / 'eig_sanitize')
    subprocess.run(['g++', '-O1', '-g', '-std=c++17', '-fno-omit-frame-pointer', '-fsanitize=address,undefined',
                    '-fno-sanitize-recover=all', '-o', exe, os.path.join(HERE, 'native', 'eig_sanitize_main.cpp')],
                   check=True)
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=300)
    out = _clean(r)
    assert '0 bad' in out
