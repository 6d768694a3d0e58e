#!/bin/bash
# Compiles the reference PLSQR3 (the C sources under /root/reference/source,
# read in place, never copied into the repository) as its FP64 configuration
# into oracle/_ref/plsqr3_ref_fp64.  Test infrastructure: the binary writes the
# golden fixtures under tests/golden/ and can be timed as the CPU baseline.
#
# main.h:26 hard-codes "#define SDVERSION 1" (float); the double build is the
# same header with SDVERSION 2 (main.h:27-35).  The recipe writes that one-line
# variant of main.h to a temporary directory, compiles the reference sources
# with the current directory dropped from the quote-include search (-I-) so the
# variant is the main.h they see, and deletes the temporary directory.
#
# Needs MPICH (the reference's MPI; /opt/conda ships MPICH 3.3.2 with mpicc).
set -euo pipefail
REF=${PLSQR_REF_SRC:-/root/reference/source}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MPICC=${MPICC:-/opt/conda/bin/mpicc}
if [ ! -f "$REF/lsqr.c" ]; then
    echo "reference sources not found at $REF; nothing built" >&2
    exit 0
fi
if [ ! -x "$MPICC" ]; then
    echo "mpicc not found at $MPICC; reference not built" >&2
    exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
sed 's/^#define SDVERSION 1 /#define SDVERSION 2 /' "$REF/main.h" > "$TMP/main.h"
grep -q '^#define SDVERSION 2 ' "$TMP/main.h"
export MPICH_CC=gcc
FLAGS="-O3 -w"
for f in main lsqr mycblas FileIO aprod loadbalance plsqr3 CPUSpMV tridiagonal vecops \
         aprod_impl vector_comm sparse_vec; do
    "$MPICC" $FLAGS -I"$TMP" -I"$REF" -I"$REF/tool" -I- -c "$REF/$f.c" -o "$TMP/$f.o" 2>/dev/null
done
for f in gptl threadutil GPTLutil; do
    "$MPICC" $FLAGS -DHAVE_MPI -DHAVE_GETTIMEOFDAY -I"$REF/tool" -I- -c "$REF/tool/$f.c" \
        -o "$TMP/$f.o" 2>/dev/null
done
"$MPICC" -o "$OUT/plsqr3_ref_fp64" "$TMP"/*.o -lm
echo "built $OUT/plsqr3_ref_fp64"
# The drop-in: the same reference objects with source/aprod.c replaced by the
# binding a maintainer would add (plsqr3.2013.09_amd/integration/aprod_plsqr3.c),
# linked against libplsqr_hip.so (built first by plsqr3.2013.09_amd/build.py).
PKG="$(dirname "$HERE")/plsqr3.2013.09_amd"
if [ -f "$PKG/libplsqr_hip.so" ]; then
    "$MPICC" $FLAGS -I"$TMP" -I"$REF" -I"$REF/tool" -I"$(dirname "$HERE")/include" -I- \
        -c "$PKG/integration/aprod_plsqr3.c" -o "$TMP/aprod_plsqr3.o"
    # linked by gcc itself so that the system libstdc++ (which the HIP runtime
    # needs) comes before the older one next to MPICH in the run-time search
    MPILIB=$(dirname "$MPICC")/../lib
    gcc -o "$OUT/plsqr3_hip_dropin" $(ls "$TMP"/*.o | grep -v '/aprod\.o$') \
        -L"$PKG" -lplsqr_hip -Wl,-rpath,'$ORIGIN/../../plsqr3.2013.09_amd' \
        -Wl,-rpath,/usr/lib/x86_64-linux-gnu -Wl,-rpath,/opt/rocm/lib -Wl,-rpath-link,/opt/rocm/lib \
        -L"$MPILIB" -lmpi -Wl,-rpath,"$MPILIB" -lm
    echo "built $OUT/plsqr3_hip_dropin"
else
    echo "libplsqr_hip.so not built yet; drop-in binary skipped" >&2
fi
# the kernel-format tools (PLSQR3_tools/kernel_format/makefile builds them the
# same way: each .c with FileIO.c where it is used); they write the golden
# digests of the example dataset's conversion chain
TOOLS=${PLSQR_REF_TOOLS:-$(dirname "$REF")/PLSQR3_tools/kernel_format}
if [ -f "$TOOLS/ker_ascii2bin.c" ]; then
    "$MPICC" $FLAGS -o "$OUT/ker_ascii2bin" "$TOOLS/ker_ascii2bin.c" 2>/dev/null
    for t in Ker2PLSQR3 Ker2PLSQR3_preprocess; do
        "$MPICC" $FLAGS -I"$REF" -o "$OUT/$t" "$TOOLS/$t.c" "$REF/FileIO.c" -lm 2>/dev/null
    done
    echo "built $OUT/ker_ascii2bin $OUT/Ker2PLSQR3 $OUT/Ker2PLSQR3_preprocess"
fi
