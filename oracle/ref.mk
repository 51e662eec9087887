# Builds the reference's own Fortran hot-path sources, where they lie under
# /root/reference, into oracle/_ref/ (test infrastructure; never shipped in the
# product).  The classic and SharpClaw routines go into separate libraries
# because both define `flux2`.  The C shim supplies the Riemann callbacks and the
# user `evec` the reference routines take (see ref_shim.c).
REF ?= /root/reference/src/pyclaw
FC = /opt/rocm/bin/amdflang
CC ?= gcc
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
FFLAGS = -O2 -fPIC -ffp-contract=off
CFLAGS = -O2 -fPIC -ffp-contract=off -std=c11

CLASSIC = $(REF)/classic/step2.f90 $(REF)/classic/flux2.f90 $(REF)/classic/limiter.f90 \
          $(REF)/classic/philim.f90 $(REF)/classic/step2ds.f90 $(REF)/classic/step3ds.f90 \
          $(REF)/classic/flux3.f90 $(REF)/classic/step3.f90
SHARP = $(REF)/sharpclaw/ClawParams.f90 $(REF)/sharpclaw/workspace.f90 \
        $(REF)/sharpclaw/weno.f90 $(REF)/sharpclaw/reconstruct.f90 \
        $(REF)/sharpclaw/flux1.f90 $(REF)/sharpclaw/flux2.f90

EX ?= /root/reference/examples
VCADV = $(EX)/advection_reaction_2d/rpn2_vc_advection_vector.f90
DEV ?= /root/reference/development/rp_approaches
EULER5 = $(DEV)/rpn2_euler_5wave.f $(DEV)/rpt2_euler_5wave.f

all: $(OUT)/libref_classic.so $(OUT)/libref_sharp.so $(OUT)/libref_vcadv.so \
     $(OUT)/libref_euler5.so

# the reference tree's euler_5wave_2D solvers (rpn2 and rpt2, older argument
# lists) with a C adapter to the lists step2/step2ds pass, in a library of their own
$(OUT)/libref_euler5.so: $(EULER5) $(HERE)ref_euler5.c
	mkdir -p $(OUT)/euler5
	cd $(OUT)/euler5 && $(FC) $(FFLAGS) -c $(EULER5)
	$(CC) $(CFLAGS) -c $(HERE)ref_euler5.c -o $(OUT)/euler5/ref_euler5.o
	$(FC) -shared -fPIC -Wl,-Bsymbolic -o $@ $(OUT)/euler5/*.o

# the advection_reaction_2d example's own Riemann solver (rpn2 only), in a
# library of its own: its symbol is rpn2, and the step takes it as an argument
$(OUT)/libref_vcadv.so: $(VCADV)
	mkdir -p $(OUT)/vcadv
	cd $(OUT)/vcadv && $(FC) $(FFLAGS) -c $(VCADV)
	$(FC) -shared -fPIC -o $@ $(OUT)/vcadv/*.o

$(OUT)/libref_classic.so: $(CLASSIC) $(HERE)ref_shim.c $(HERE)oracle.c $(HERE)oracle.h
	mkdir -p $(OUT)/classic
	cd $(OUT)/classic && $(FC) $(FFLAGS) -c $(CLASSIC)
	$(CC) $(CFLAGS) -DREF_CLASSIC -c $(HERE)ref_shim.c -o $(OUT)/classic/ref_shim.o
	$(CC) $(CFLAGS) -c $(HERE)oracle.c -o $(OUT)/classic/oracle.o
	$(FC) -shared -fPIC -Wl,-Bsymbolic -o $@ $(OUT)/classic/*.o -lm

$(OUT)/libref_sharp.so: $(SHARP) $(HERE)ref_sharp_driver.f90 $(HERE)ref_shim.c $(HERE)oracle.c $(HERE)oracle.h
	mkdir -p $(OUT)/sharp
	cd $(OUT)/sharp && $(FC) $(FFLAGS) -c $(SHARP) $(HERE)ref_sharp_driver.f90
	$(CC) $(CFLAGS) -DREF_SHARP -c $(HERE)ref_shim.c -o $(OUT)/sharp/ref_shim.o
	$(CC) $(CFLAGS) -c $(HERE)oracle.c -o $(OUT)/sharp/oracle.o
	$(FC) -shared -fPIC -Wl,-Bsymbolic -o $@ $(OUT)/sharp/*.o -lm

clean:
	rm -rf $(OUT)
