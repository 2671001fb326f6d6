#!/usr/bin/env bash
# Compiles the reference routines that build from their own sources without FFTW
# (the reference's only external dependency) into oracle/_ref/libref.so, behind the
# bind(C) shim oracle/ref_shim.f90.  Sources are compiled where they lie in the
# reference tree; nothing is copied.  Test infrastructure only.
set -euo pipefail
REF=${REF:-/root/reference/Bulk_3D_FFT}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
OUT="$HERE/_ref"
FLANG=${FLANG:-/opt/rocm/llvm/bin/flang}
if [ ! -d "$REF" ] || [ ! -x "$FLANG" ]; then
  echo "build_ref.sh: reference tree or flang missing; oracle/_ref not built" >&2
  exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
for f in "$REF/Main_Diblock_Melt/mod_global.f90" "$REF/General_Routines/mod_mt.f90" \
         "$REF/General_Routines/mod_iterate.f90" "$REF/Interaction_Modules/mod_FloryHuggins.f90" \
         "$HERE/ref_shim.f90"; do
  "$FLANG" -O2 -fPIC -c "$f"
done
"$FLANG" -shared -fPIC -o libref.so mod_global.o mod_mt.o mod_iterate.o mod_FloryHuggins.o ref_shim.o
echo "built $OUT/libref.so"
