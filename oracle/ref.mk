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

all: $(OUT)/thorin

$(OUT):
	mkdir -p $(OUT)

$(OUT)/main_%.o: $(REF)/src_main/%.c | $(OUT)
	$(CC) -c $(MAIN_CFLAGS) -o $@ $<

$(OUT)/reb_%.o: $(REF)/src_reb/%.c | $(OUT)
	$(CC) -c $(REB_CFLAGS) -o $@ $<

$(OUT)/thorin: $(MAIN_OBJ) $(REB_OBJ)
	$(CC) -O3 -o $@ $(MAIN_OBJ) $(REB_OBJ) -lm -lrt

clean:
	rm -rf $(OUT)

.PHONY: all clean
