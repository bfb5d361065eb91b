"""CPU checks of the C-ABI boundary: the built library loads and exports every
symbol the public headers declare.  No device code runs here."""
import os
import re
import subprocess

import pytest

import halloc_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NM = "nm"


def declared_c_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(ha_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(H.LIB_PATH):
        from halloc_amd import build
        build.build()
    return H.LIB_PATH


def test_c_abi_exports_every_declared_function(built):
    names = declared_c_functions("halloc_amd.h")
    assert len(names) >= 14
    L = H.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(H.EXPORTS)


def test_drop_in_host_api_exported(built):
    # ha_init(halloc_opts_t), ha_shutdown(), ha_extfrag(bool) of include/halloc.h
    out = subprocess.run([NM, "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    for mangled in ("_Z7ha_init13halloc_opts_t", "_Z11ha_shutdownv", "_Z10ha_extfragb"):
        assert mangled in out, mangled


LLVM = "/opt/rocm/lib/llvm/bin"


def test_rdc_archive_holds_device_api(built, tmp_path):
    # libhalloc.a carries the host API and, as relocatable gfx950 device code
    # (LLVM bitcode in the object's offload bundle), the device definitions of
    # hamalloc(size_t), hafree(void*) and the batched calls that user code
    # built with -fgpu-rdc links against
    archive = os.path.join(ROOT, "halloc_amd", "lib", "libhalloc.a")
    assert os.path.exists(archive)
    out = subprocess.run([NM, archive], capture_output=True, text=True).stdout
    assert "_Z7ha_init13halloc_opts_t" in out
    subprocess.run([f"{LLVM}/llvm-ar", "x", archive, "halloc.o"], cwd=tmp_path, check=True)
    targets = subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", "--input=halloc.o", "--list"],
                             cwd=tmp_path, capture_output=True, text=True, check=True).stdout.split()
    dev = [t for t in targets if t.endswith("gfx950")]
    assert dev, targets
    subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", "--input=halloc.o", f"--targets={dev[0]}",
                    "--output=dev.bc", "--unbundle"], cwd=tmp_path, check=True)
    ir = subprocess.run([f"{LLVM}/llvm-dis", "dev.bc", "-o", "-"], cwd=tmp_path, capture_output=True, text=True,
                        check=True).stdout
    defined = set(re.findall(r"^define [^@]*@(\S+?)\(", ir, flags=re.M))
    for sym in ("_Z8hamallocm", "_Z6hafreePv", "_Z14hamalloc_batchmiPPv", "_Z12hafree_batchPKPvi"):
        assert sym in defined, sym


def test_default_options_match_reference(built):
    # src/halloc.h:30-32 defaults
    o = H._Opts()
    H.lib().ha_default_opts(o)
    assert (o.memory, o.halloc_fraction, o.busy_fraction, o.roomy_fraction, o.sparse_fraction,
            o.sb_sz_sh) == (512 << 20, 0.75, 0.835, 0.6, 0.012, 22)


def test_calls_without_init_fail_cleanly(built):
    with pytest.raises(H.HallocError, match="not initialized"):
        H.Halloc.geometry()


def test_product_never_imports_the_oracle():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "halloc_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)
