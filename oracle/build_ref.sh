#!/bin/sh
# Builds oracle/_ref/libref_fallback.so: the reference's header-only CPU fallback
# (include/parallax/stdpar.hpp of the reference checkout, used where it lies) wrapped by
# ref_fallback_shim.cpp and linked to libparallax_amd.so for the runtime.h symbols.
# Flags are the reference's Release defaults (CMakeLists.txt:4-5): -O3 -DNDEBUG -std=c++20.
# Output goes only to oracle/_ref/ (git-ignored).  Skips quietly without the reference.
set -e
REF="${PARALLAX_REFERENCE:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
LIBDIR="$HERE/../parallax-runtime_amd/lib"
if [ ! -f "$REF/include/parallax/stdpar.hpp" ]; then
  echo "build_ref: reference checkout not found at $REF; oracle/_ref not built"
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O3 -DNDEBUG -std=c++20 -fPIC -shared -I"$REF/include" "$HERE/ref_fallback_shim.cpp" \
    -L"$LIBDIR" -lparallax_amd -Wl,-rpath,'$ORIGIN/../../parallax-runtime_amd/lib' \
    -o "$HERE/_ref/libref_fallback.so"
echo "build_ref: built $HERE/_ref/libref_fallback.so"
