#!/bin/sh
# Writes siphash13_kat.txt: SipHash-1-3 (key 0) of the 8-byte little-endian
# encoding of each value, computed by OpenSSL (an implementation independent
# of this repository).  Rust's `usize: Hash` feeds exactly these 8 bytes.
set -e
out="$(dirname "$0")/siphash13_kat.txt"
: > "$out"
for v in 0 1 2 3 255 256 65535 400 16383 12345678 4294967295 4294967296 18446744073709551615; do
  hex=$(python3 -c "import struct,sys; sys.stdout.write(struct.pack('<Q', $v).hex())")
  mac=$(python3 -c "import sys; sys.stdout.buffer.write(bytes.fromhex('$hex'))" | \
        openssl mac -macopt hexkey:00000000000000000000000000000000 -macopt size:8 \
                    -macopt c-rounds:1 -macopt d-rounds:3 SIPHASH)
  echo "$v $mac" >> "$out"
done
