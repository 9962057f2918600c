"""Writes chacha_vectors.json: ChaCha20 keystreams (RFC 7539) computed by the
openssl command-line tool, used to pin the oracle's and the GPU's ChaCha20
and the PRNG stream built on it.

    python tests/golden/make_chacha_vectors.py

openssl takes a 16-byte IV: the 32-bit little-endian initial block counter
followed by the 96-bit nonce.  The PRNG streams follow the reference's
construction (prng/chacha_prng_util.cc:27-60): buffer s is the keystream of
255 * 32 bytes under nonce "salt" || LE64(s), block counter 0.
"""
import hashlib
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BUF = 255 * 32


def keystream(key, counter, nonce, nbytes):
    iv = counter.to_bytes(4, "little") + nonce
    r = subprocess.run(["openssl", "enc", "-chacha20", "-K", key.hex(), "-iv", iv.hex()],
                       input=bytes(nbytes), capture_output=True, check=True)
    assert len(r.stdout) == nbytes
    return r.stdout


def main():
    out = {"tool": subprocess.run(["openssl", "version"], capture_output=True,
                                  text=True).stdout.strip(), "blocks": [], "prng_streams": []}
    rfc_key = bytes(range(32))
    rfc_nonce = bytes.fromhex("000000090000004a00000000")
    for key, counter, nonce in [(rfc_key, 1, rfc_nonce), (bytes(32), 0, bytes(12)),
                                (bytes(range(100, 132)), 7, b"salt" + bytes(8))]:
        ks = keystream(key, counter, nonce, 64)
        out["blocks"].append({"key": key.hex(), "counter": counter, "nonce": nonce.hex(),
                              "keystream": ks.hex()})
    for key in (bytes(range(32)), hashlib.sha256(b"shell-gpu prng key").digest()):
        bufs = []
        for salt in range(3):
            ks = keystream(key, 0, b"salt" + salt.to_bytes(8, "little"), BUF)
            bufs.append({"salt": salt, "sha256": hashlib.sha256(ks).hexdigest(),
                         "head": ks[:16].hex(), "tail": ks[-16:].hex()})
        out["prng_streams"].append({"key": key.hex(), "buffers": bufs})
    with open(os.path.join(HERE, "chacha_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
