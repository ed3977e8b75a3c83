#!/usr/bin/env bash
# ref_build.sh -- compile the reference's own per-step sources, where they lie
# under $MLAU_REF (default /root/reference), into oracle/_ref/*.so.  Nothing
# from the reference is copied: each solver is streamed through sed (which
# only rewrites its scheme-selection lines, the reference's documented way of
# choosing ODR / R_K / func_flux) straight into g++ with the reference's own
# flags (Makefile.inc:1-4: g++ -O2 -fopenmp).  Outputs are test-only checkers.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${MLAU_REF:-/root/reference}"
OUT="$HERE/_ref"
OBJ="$OUT/obj"
CXX="${CXX:-g++}"
CFLAGS="-O2 -fopenmp -fPIC -w"
if [ ! -d "$REF/include" ]; then
  echo "ref_build: no reference tree at $REF, skipping" >&2
  exit 0
fi
mkdir -p "$OBJ"

$CXX $CFLAGS -I"$REF/include" -shared "$HERE/ref_kat.cpp" -o "$OUT/libref_kat.so"
$CXX $CFLAGS -I"$REF/include" -I"$REF/2D/SERIAL/OTvortex" -shared "$HERE/ref_cfl_glue.cpp" -o "$OUT/libref_cfl2d.so"

flux_fn() {
  case "$1" in
    mlau) echo calc_flux_mlau ;; roe) echo calc_flux_roe ;;
    hlld) echo calc_flux_hlld ;; lhlld) echo calc_flux_lhlld ;;
  esac
}

# select <src> <odr> <rk> <flux> <ctw>: stream a solver with its scheme lines set
select_src() {
  sed -e "s/^#define ODR ([0-9])/#define ODR ($2)/" \
      -e "s/^#define R_K ([0-9])/#define R_K ($3)/" \
      -e "s/^#define CTW ([0-9])/#define CTW ($5)/" \
      -e "s/^\(  *\)func_flux=&calc_flux_[a-z0-9]*;/\1func_flux=\&$(flux_fn "$4");/" "$1"
}

build_2d() {  # name odr rk flux ctw
  local d="$REF/2D/SERIAL/OTvortex"
  select_src "$d/mhd_fd_ct_2d.c" "$2" "$3" "$4" "$5" |
    $CXX $CFLAGS -I"$REF/include" -x c++ -c - -o "$OBJ/2d_$1.o"
  $CXX $CFLAGS -x c++ -c "$d/boundary.c" -o "$OBJ/2d_bc.o"
  $CXX $CFLAGS -DREF_DIM=2 -c "$HERE/ref_step_glue.cpp" -o "$OBJ/glue2.o"
  $CXX $CFLAGS -shared "$OBJ/2d_$1.o" "$OBJ/2d_bc.o" "$OBJ/glue2.o" -o "$OUT/libref2d_$1.so"
}

build_1d() {  # name odr rk flux kind(shock|wave)
  local d="$REF/1D/SERIAL/$5"
  select_src "$d/mhd_fd4c_1d.c" "$2" "$3" "$4" 1 |
    $CXX $CFLAGS -I"$REF/include" -x c++ -c - -o "$OBJ/1d_$1.o"
  if [ "$5" = shock ]; then
    $CXX $CFLAGS -I"$REF/include" -x c++ -c "$d/openbc1d_mhd.c" -o "$OBJ/1d_bc_$5.o"
  else
    $CXX $CFLAGS -x c++ -c "$d/prdc_1d.c" -o "$OBJ/1d_bc_$5.o"
  fi
  $CXX $CFLAGS -DREF_DIM=1 -c "$HERE/ref_step_glue.cpp" -o "$OBJ/glue1.o"
  $CXX $CFLAGS -shared "$OBJ/1d_$1.o" "$OBJ/1d_bc_$5.o" "$OBJ/glue1.o" -o "$OUT/libref1d_$1.so"
}

MPI_INC="${MPI_INC:-/opt/conda/include}"
MPI_LIB="${MPI_LIB:-/opt/conda/lib}"
build_3d() {  # name odr rk flux ctw
  local d="$REF/3D/MPI/OTvortex"
  if [ ! -f "$MPI_INC/mpi.h" ]; then
    echo "ref_build: no mpi.h at $MPI_INC; 3D reference build skipped" >&2
    return 0
  fi
  select_src "$d/mhd_fd_ct_3d.c" "$2" "$3" "$4" "$5" |
    $CXX $CFLAGS -I"$REF/include" -x c++ -c - -o "$OBJ/3d_$1.o"
  $CXX $CFLAGS -I"$MPI_INC" -x c++ -c "$d/boundary.c" -o "$OBJ/3d_bc.o"
  $CXX $CFLAGS -DREF_DIM=3 -c "$HERE/ref_step_glue.cpp" -o "$OBJ/glue3.o"
  $CXX $CFLAGS -shared "$OBJ/3d_$1.o" "$OBJ/3d_bc.o" "$OBJ/glue3.o" -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpi \
    -o "$OUT/libref3d_$1.so"
}

# 2D variants: name odr rk flux ctw
build_2d o4r3_lhlld 4 3 lhlld 1
build_2d o2r3_mlau 2 3 mlau 1
build_2d o2r3_hlld 2 3 hlld 1
build_2d o2r3_roe 2 3 roe 1
build_2d o2r3_lhlld 2 3 lhlld 1
build_2d o1r3_lhlld 1 3 lhlld 1
build_2d o3r3_mlau 3 3 mlau 1
build_2d o4r2_mlau 4 2 mlau 1
build_2d o2r1_hlld 2 1 hlld 1
build_2d o2r3_lhlld_ctw0 2 3 lhlld 0
# 1D variants: name odr rk flux kind
build_1d shock_o2r3_mlau 2 3 mlau shock
build_1d shock_o1r2_lhlld 1 2 lhlld shock
build_1d shock_o4r3_lhlld 4 3 lhlld shock
build_1d shock_o3r3_hlld 3 3 hlld shock
build_1d shock_o2r3_roe 2 3 roe shock
build_1d wave_o2r3_lhlld 2 3 lhlld wave
build_1d wave_o4r3_mlau 4 3 mlau wave
# 3D variants
if [ -f "$MPI_INC/mpi.h" ]; then
  $CXX $CFLAGS -I"$REF/include" -I"$REF/3D/MPI/OTvortex" -I"$MPI_INC" -shared "$HERE/ref_cfl3d_glue.cpp" \
    -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpi -o "$OUT/libref_cfl3d.so"
fi
build_3d o4r3_mlau 4 3 mlau 1
build_3d o2r3_lhlld 2 3 lhlld 1
build_3d o2r3_hlld 2 3 hlld 1
build_3d o3r3_roe 3 3 roe 1
# The reference's own 2D driver (2D/SERIAL/OTvortex/main.cpp with boundary.c,
# unmodified: 200^2 Orszag-Tang vortex, 750 steps, outdat every 150), linked
# once with the reference solver and once with this project's GPU drop-in
# library (libmlau_dropin, built beforehand in mlau_amd/_lib) in place of
# mhd_fd_ct_2d.o.  The test suite runs both and compares their dat/ files.
DROPIN="$HERE/../mlau_amd/_lib"
OTV="$REF/2D/SERIAL/OTvortex"
$CXX $CFLAGS -I"$REF/include" -c "$OTV/main.cpp" -o "$OBJ/otv_main.o"
$CXX $CFLAGS -x c++ -c "$OTV/boundary.c" -o "$OBJ/otv_bc.o"
$CXX $CFLAGS -I"$REF/include" -x c++ -c "$OTV/mhd_fd_ct_2d.c" -o "$OBJ/otv_solver.o"
$CXX $CFLAGS "$OBJ/otv_main.o" "$OBJ/otv_solver.o" "$OBJ/otv_bc.o" -o "$OUT/otvortex2d_ref" -lm
if [ -f "$DROPIN/libmlau_dropin.so" ]; then
  $CXX $CFLAGS "$OBJ/otv_main.o" "$OBJ/otv_bc.o" -L"$DROPIN" -lmlau_dropin -lmlau_hip \
    -Wl,-rpath,'$ORIGIN/../../mlau_amd/_lib' -o "$OUT/otvortex2d_gpu" -lm
else
  echo "ref_build: no libmlau_dropin.so in $DROPIN; the GPU-linked driver is not built" >&2
fi
# The MPI drivers' merge tools (per-rank FP32 outdat -> merge_outdat), to
# check that this library's output files load in the reference's own tools.
gcc -O2 -w "$REF/2D/MPI/common/merge.c" -o "$OUT/merge2d"
gcc -O2 -w "$REF/3D/MPI/common/merge.c" -o "$OUT/merge3d"
rm -rf "$OBJ"
echo "ref_build: reference checkers in $OUT"
