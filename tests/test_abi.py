"""The C ABI library loads on a machine without a GPU and exports every
function include/rsieve.h declares; the command-line host starts."""
import ctypes
import os
import re
import subprocess

from conftest import ROOT

PKG = os.path.join(ROOT, "cuda-riesel-sieve_amd")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rsieve.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("rs_abcd_read", "rs_plan_create", "rs_primes_batch", "rs_sieve_create",
              "rs_sieve_submit", "rs_sieve_collect", "rs_sieve_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(rs):
    lib = ctypes.CDLL(os.path.join(PKG, "librsieve.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol(rs):
    assert sorted(rs._SIGS) == declared_functions()


def declared_arities():
    """{function: number of parameters} from the header's prototypes."""
    text = open(os.path.join(ROOT, "include", "rsieve.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(rs_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_binding_arity_matches_header(rs):
    # a ctypes argtypes list of the wrong length would pass garbage silently
    ar = declared_arities()
    bad = {n: (len(rs._SIGS[n][1]), ar[n]) for n in ar if len(rs._SIGS[n][1]) != ar[n]}
    assert not bad, bad


def test_no_device_is_not_fatal(rs):
    # device_count() must answer (0 here) without aborting the process
    assert rs.device_count() >= 0
    assert rs.strerror(rs.RS_ERR_FORMAT) == "malformed ABCD input"


def test_cli_help_and_bad_flags():
    exe = os.path.join(PKG, "rsieve")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "-Q Q" in r.stdout
    r = subprocess.run([exe, "-Q", "3"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "Q must be an even integer" in r.stdout
    r = subprocess.run([exe, "-p", "100", "-P", "50"], capture_output=True, text=True,
                       timeout=30)
    assert r.returncode == 1 and "P1 < P0" in r.stdout
    r = subprocess.run([exe, "-m", "16"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "-m must be in [0, 15]" in r.stdout
    r = subprocess.run([exe, "-i", "/nonexistent.abcd", "-p", "1000", "-P", "2000"],
                       capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "Unable to open file" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    why, before touching a device."""
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
