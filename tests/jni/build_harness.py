"""Builds tests/jni/libjni_harness.so: the JNI shim (cartogenesis_amd/jni/cgocean_jni.c) compiled against
the test-only jni.h beside this file, with the fake JVM environment, linked to libcgocean.so.
Test infrastructure: only the tests load it."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "libjni_harness.so"


def build() -> Path:
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-Werror", f"-I{HERE}",
           f"-I{ROOT / 'include'}", str(ROOT / "cartogenesis_amd" / "jni" / "cgocean_jni.c"), str(HERE / "fake_jvm.c"),
           f"-L{ROOT / 'cartogenesis_amd'}", "-lcgocean", "-Wl,-rpath,$ORIGIN/../../cartogenesis_amd",
           "-o", str(OUT)]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
