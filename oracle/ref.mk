# Builds the reference FARGO_THORIN (sequential build, mpi_dummy) from its own
# sources under /root/reference into oracle/_ref/.  Nothing is written outside
# oracle/_ref/.  Flags follow src_main/makefile (FARGO_ARCH=THORIN) and
# src_main/makefile.reb, except that -march=native is dropped so that the
# binary also runs on another host CPU: with -std=c99 (no FMA contraction) and
# without -ffast-math this does not change any floating-point result.
#
#   make -f oracle/ref.mk            (from the repository root)
REF      ?= /root/reference
OUT      := oracle/_ref
CC       := gcc
MAIN_CFLAGS := -O3 -Wall -Wno-unknown-pragmas -Wno-unused-result -std=c99 -Wpointer-arith \
               -I$(REF)/src_reb -I$(REF)/src_main
REB_CFLAGS  := -std=c99 -Wpointer-arith -Wno-unused-result -D_GNU_SOURCE -O3 -fPIC -DLIBREBOUND -w

MAIN_SRC := LowTasks SideEuler Output Init main Theo Interpret SourceEuler TransportEuler \
            Planet Viscosity Psys Force Drag var Pframeforce split merge commbound fpe rebin \
            mpiTasks EnergySources Opacity Opacity2 ReboundInterface Pebbles mpi_dummy
REB_SRC  := rebound tree particle gravity integrator integrator_whfast integrator_whfasthelio \
            integrator_ias15 integrator_sei integrator_leapfrog integrator_hermes boundary input \
            output collision communication_mpi display tools derivatives simulationarchive glad

MAIN_OBJ := $(MAIN_SRC:%=$(OUT)/main_%.o)
REB_OBJ  := $(REB_SRC:%=$(OUT)/reb_%.o)

# The reference's two parallel builds, for CPU timing only (not parity
# oracles: the OpenMP red-black SOR colours race, EnergySources.c:459-484, and
# the MPI build splits the ring sums over ranks):
#   thorin_omp : FARGO_ARCH=THORIN_OPENMP, BUILD=sequential (make_omp.sh,
#                makefile:39-46): -fopenmp -DOPENMP, REBOUND with OPENMP=1
#   thorin_mpi : FARGO_ARCH=THORIN, BUILD=parallel (makefile:32-37, 109-113):
#                mpicc -D_PARALLEL without mpi_dummy (MPICH wrapper, host gcc);
#                built only where an
#                mpicc exists (MPICC, else mpicc on PATH, else /opt/conda/bin)
MPICC ?= $(shell command -v mpicc 2>/dev/null || (test -x /opt/conda/bin/mpicc && echo /opt/conda/bin/mpicc))
OMP_OBJ  := $(MAIN_SRC:%=$(OUT)/omp_main_%.o) $(REB_SRC:%=$(OUT)/omp_reb_%.o)
MPI_MAIN := $(filter-out mpi_dummy,$(MAIN_SRC))
MPI_OBJ  := $(MPI_MAIN:%=$(OUT)/mpi_main_%.o) $(REB_OBJ)

# The reference's own driver with this build's AlgoGas (the drop-in, for the
# tests): its objects as above, SourceEuler.o's AlgoGas made weak so that the
# shim fargo_thorin_amd/shim/algogas_shim.c (compiled with the reference's
# headers) replaces it, linked with fargo_thorin_amd/libthorin_amd.so.
SHIM    := fargo_thorin_amd/shim/algogas_shim.c
LIBGPU  := fargo_thorin_amd/libthorin_amd.so
GPU_OBJ := $(filter-out $(OUT)/main_SourceEuler.o,$(MAIN_OBJ)) $(OUT)/main_SourceEuler_weak.o \
           $(OUT)/shim_algogas.o $(REB_OBJ)

all: $(OUT)/thorin $(OUT)/thorin_omp $(if $(MPICC),$(OUT)/thorin_mpi) $(if $(wildcard $(LIBGPU)),$(OUT)/thorin_gpu)

$(OUT):
	mkdir -p $(OUT)

$(OUT)/main_%.o: $(REF)/src_main/%.c | $(OUT)
	$(CC) -c $(MAIN_CFLAGS) -o $@ $<

$(OUT)/reb_%.o: $(REF)/src_reb/%.c | $(OUT)
	$(CC) -c $(REB_CFLAGS) -o $@ $<

$(OUT)/thorin: $(MAIN_OBJ) $(REB_OBJ)
	$(CC) -O3 -o $@ $(MAIN_OBJ) $(REB_OBJ) -lm -lrt

$(OUT)/omp_main_%.o: $(REF)/src_main/%.c | $(OUT)
	$(CC) -c $(MAIN_CFLAGS) -fopenmp -DOPENMP -o $@ $<

$(OUT)/omp_reb_%.o: $(REF)/src_reb/%.c | $(OUT)
	$(CC) -c $(REB_CFLAGS) -fopenmp -DOPENMP -o $@ $<

$(OUT)/thorin_omp: $(OMP_OBJ)
	$(CC) -O3 -fopenmp -o $@ $(OMP_OBJ) -lm -lrt

$(OUT)/mpi_main_%.o: $(REF)/src_main/%.c | $(OUT)
	$(MPICC) -cc=$(CC) -c $(MAIN_CFLAGS) -D_PARALLEL -o $@ $<

$(OUT)/thorin_mpi: $(MPI_OBJ)
	$(MPICC) -cc=$(CC) -O3 -o $@ $(MPI_OBJ) -lm -lrt

clean:
	rm -rf $(OUT)

.PHONY: all clean

$(OUT)/main_SourceEuler_weak.o: $(OUT)/main_SourceEuler.o
	objcopy --weaken-symbol=AlgoGas $< $@

$(OUT)/shim_algogas.o: $(SHIM) include/thorin_amd.h | $(OUT)
	$(CC) -c $(MAIN_CFLAGS) -Iinclude -o $@ $<

$(OUT)/thorin_gpu: $(GPU_OBJ) $(LIBGPU)
	$(CC) -O3 -o $@ $(GPU_OBJ) -Lfargo_thorin_amd -lthorin_amd -Wl,-rpath,'$$ORIGIN/../../fargo_thorin_amd' -lm -lrt
