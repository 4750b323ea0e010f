"""Writes ``save_gzip_64x32.cgwd``: the reference's own whole-save fixture, as the bytes it encodes.

The reference's tests hold a complete format-28 save of a 64 x 32 world (seed 4096, gzip chunks) as
base64 text split over string literals (cartography/src/commonTest/kotlin/com/cartogenesis/
cartography/GzipFixture.kt, made by RegenerateGzipFixture.kt:18-30). This script reads that file as
text, joins the literals, decodes the base64 and writes the file's bytes here: data, not source.
Run it in a checkout that has the reference at /root/reference; the output is committed.
"""
import base64
import hashlib
import re
from pathlib import Path

SOURCE = Path("/root/reference/cartography/src/commonTest/kotlin/com/cartogenesis/cartography/GzipFixture.kt")
OUT = Path(__file__).resolve().parent / "save_gzip_64x32.cgwd"


def main() -> None:
    text = SOURCE.read_text()
    body = text.split("GZIP_FIXTURE_BASE64: String =", 1)[1]
    data = base64.b64decode("".join(re.findall(r'"([^"]*)"', body)))
    assert data[:4] == b"CGWD", "not a save"
    OUT.write_bytes(data)
    print(f"wrote {OUT.name}: {len(data)} bytes, sha256 {hashlib.sha256(data).hexdigest()}")


if __name__ == "__main__":
    main()
