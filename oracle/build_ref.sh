#!/bin/bash
# Builds the reference's CPU entropy coder (MLCodec_extensions_cpp) from its own
# sources under /root/reference into oracle/_ref/, for use as a test oracle only.
# The reference's setup.py requires Python >= 3.12, so g++ is driven directly
# with the same flags (src/cpp/setup.py:14).  Output stays in oracle/_ref/.
set -e
REF=${DCVC_REFERENCE:-/root/reference}
SRC=$REF/src/cpp/py_rans
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
if [ ! -d "$SRC" ]; then echo "reference sources not present; skipping"; exit 0; fi
mkdir -p "$OUT"
SUFFIX=$(python3 -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
PYINC=$(python3 -c "import sysconfig; print(sysconfig.get_paths()['include'])")
PBINC=$(python3 -c "import pybind11; print(pybind11.get_include())")
g++ -std=c++17 -O3 -fPIC -shared -Wall -Wextra -I"$PYINC" -I"$PBINC" \
    "$SRC/rans.cpp" "$SRC/py_rans.cpp" "$SRC/bind.cpp" \
    -o "$OUT/MLCodec_extensions_cpp$SUFFIX" -lpthread
echo "built $OUT/MLCodec_extensions_cpp$SUFFIX"
