#!/usr/bin/env bash
# Builds the reference engine (asavine/CompFinance, header-only C++17) with the
# driver ref_driver.cpp into oracle/_ref/ref_driver.  The reference is read
# from REF_DIR (default /root/reference); nothing of it is kept in the
# repository.  Three mechanical portability fixes are applied to a throw-away
# copy in a temporary directory (deleted on exit):
#   1. the reference includes "ThreadPool.h" but ships threadPool.h;
#   2. AADExpr.h calls the member template pushAdjoint on a dependent object
#      without the `template` keyword, which clang rejects;
#   3. mrg32k3a.h uses MSVC-only functional casts `unsigned long long(x)`.
# The Excel front end (xl*) is not compiled.
set -euo pipefail
REF_DIR=${REF_DIR:-/root/reference}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
OUT=$HERE/_ref
if [ ! -f "$REF_DIR/mcBase.h" ]; then
  echo "reference sources not found in $REF_DIR; skipping reference build" >&2
  exit 0
fi
CXX=${REF_CXX:-}
if [ -z "$CXX" ]; then
  if [ -x /opt/rocm/lib/llvm/bin/clang++ ]; then CXX=/opt/rocm/lib/llvm/bin/clang++; else CXX=clang++; fi
fi
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
for f in "$REF_DIR"/*.h "$REF_DIR"/*.cpp; do
  case "$(basename "$f")" in xl*) continue ;; esac
  cp "$f" "$TMP/"
done
cp "$TMP/threadPool.h" "$TMP/ThreadPool.h"
sed -i 's/\.pushAdjoint</.template pushAdjoint</g' "$TMP/AADExpr.h"
sed -i -E 's/unsigned long long ?\(([^()]*)\)/((unsigned long long)(\1))/g' "$TMP/mrg32k3a.h"
mkdir -p "$OUT"
"$CXX" -std=c++17 -O3 -w -I"$TMP" -include cstring -include algorithm -include functional \
  "$HERE/ref_driver.cpp" "$TMP/sobol.cpp" "$TMP/AAD.cpp" "$TMP/mcBase.cpp" "$TMP/ThreadPool.cpp" \
  -pthread -o "$OUT/ref_driver"
echo "built $OUT/ref_driver with $CXX"
