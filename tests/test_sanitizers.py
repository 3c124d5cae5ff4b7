"""CPU: the library's host side and the C restatement of galario's CPU loops under AddressSanitizer
and UndefinedBehaviorSanitizer; no GPU is needed.

* Host glue: galario_amd.hip built with the sanitizers on its host compilation only (the kernels are
  compiled as usual; about a minute) and linked into tests/sanitize/host_glue.c, which drives the
  argument checks,
  error strings and plan handling of the C ABI: every call must fail cleanly with a status and a
  message, and neither sanitizer may report.
* The C port: oracle/galario_port.c linked into tests/sanitize/port.c, which runs each of its
  functions once on a small problem with consistent buffers.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, 'include')
CSRC = os.path.join(ROOT, 'galario_amd', 'csrc')
ROCM = '/opt/rocm'
SAN_ENV = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:halt_on_error=1',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1')


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, **kw)


def _clean(r):
    out = r.stdout + r.stderr
    return r.returncode == 0 and 'ALL OK' in r.stdout and 'Sanitizer' not in out and 'runtime error' not in out


def test_host_glue_under_sanitizers(tmp_path):
    hipcc = shutil.which('hipcc') or os.path.join(ROCM, 'bin', 'hipcc')
    clang = os.path.join(ROCM, 'llvm', 'bin', 'clang')
    if not (os.path.exists(hipcc) and os.path.exists(clang)):
        pytest.skip('hipcc or clang not available')
    lib = str(tmp_path / 'libgalario_amd_san.so')
    r = _run([hipcc, '--offload-arch=gfx950', '-O1', '-std=c++17', '-fPIC', '-shared', '-fno-omit-frame-pointer',
              '-Xarch_host', '-g', '-Xarch_host', '-fsanitize=address', '-Xarch_host', '-fsanitize=undefined', '-w',
              '-I', INCLUDE,
              '-I', CSRC, os.path.join(CSRC, 'galario_amd.hip'), '-o', lib], timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = str(tmp_path / 'host_glue')
    r = _run([clang, '-O1', '-g', '-std=c11', '-fno-omit-frame-pointer', '-fsanitize=address,undefined', '-I', INCLUDE,
              os.path.join(HERE, 'sanitize', 'host_glue.c'), lib, '-L' + os.path.join(ROCM, 'lib'), '-lamdhip64',
              '-Wl,-rpath,' + str(tmp_path) + ':' + os.path.join(ROCM, 'lib'), '-o', exe], timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    r = _run([exe], env=SAN_ENV, timeout=300)
    print(r.stdout[-4000:])
    assert _clean(r), (r.returncode, r.stdout[-3000:], r.stderr[-3000:])


def test_c_port_under_sanitizers(tmp_path):
    gcc = shutil.which('gcc')
    if not gcc:
        pytest.skip('gcc not available')
    exe = str(tmp_path / 'port')
    r = _run([gcc, '-O1', '-g', '-std=c11', '-D_DEFAULT_SOURCE', '-fopenmp', '-ffp-contract=off',
              '-fno-omit-frame-pointer', '-fsanitize=address,undefined', os.path.join(ROOT, 'oracle', 'galario_port.c'),
              os.path.join(HERE, 'sanitize', 'port.c'), '-lm', '-o', exe], timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    r = _run([exe], env=dict(SAN_ENV, OMP_NUM_THREADS='2'), timeout=300)
    print(r.stdout[-2000:])
    assert _clean(r), (r.returncode, r.stdout[-3000:], r.stderr[-3000:])
