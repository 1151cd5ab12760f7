# Builds the unmodified MOLSIM reference (serial, without FFTW) from the
# Fortran sources where they lie, into oracle/_ref/.  Test infrastructure
# only: the binary is run by tests/golden/make_golden.py to produce the
# committed golden fixtures; nothing in the product links or runs it.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
#
# Flags follow the reference's LOCAL_GFORTRAN "normal" preset
# (Src/makefile:160-190: -ffree-form -O2), with amdflang as the compiler.
# F03_CBIND is left undefined, so SPME (which needs FFTW3, absent here) is
# compiled out exactly as Src/energy.F90:6857-6858 arranges; _PAR_ is
# undefined (no MPI).

REF    ?= /root/reference
SRC    := $(REF)/Src
OUT    := $(dir $(lastword $(MAKEFILE_LIST)))_ref
REF_FC ?= amdflang
FFLAGS ?= -cpp -ffree-form -O2

# Module dependency order (Src/makefile:406-460).
UNITS := mollib mesh statistics mol molaux celllist particle potential \
         coordinate md bd nlist energy denergy mc dump group static dynamic \
         image mixed moluser sso
OBJS  := $(addprefix $(OUT)/,$(addsuffix .o,$(UNITS)))

all: $(OUT)/molsim $(OUT)/ref_probe

$(OUT):
	mkdir -p $(OUT)

$(OUT)/%.o: $(SRC)/%.F90 | $(OUT)
	cd $(OUT) && $(REF_FC) $(FFLAGS) -c $< -o $(notdir $@)

$(OUT)/mol.o: FFLAGS += -DFLIBMACRO=\"$(SRC)/molsim.lib\"

# serialised module chain: each unit needs the .mod files of its predecessors
$(OUT)/mesh.o: $(OUT)/mollib.o
$(OUT)/statistics.o: $(OUT)/mollib.o
$(OUT)/mol.o: $(OUT)/statistics.o $(OUT)/mesh.o $(OUT)/mollib.o
$(OUT)/molaux.o $(OUT)/celllist.o $(OUT)/particle.o $(OUT)/potential.o \
$(OUT)/coordinate.o $(OUT)/md.o $(OUT)/dump.o $(OUT)/group.o \
$(OUT)/image.o $(OUT)/mixed.o: $(OUT)/mol.o
$(OUT)/bd.o: $(OUT)/md.o
$(OUT)/nlist.o $(OUT)/energy.o: $(OUT)/celllist.o
$(OUT)/denergy.o: $(OUT)/energy.o
$(OUT)/mc.o: $(OUT)/nlist.o
$(OUT)/static.o $(OUT)/moluser.o: $(OUT)/potential.o
$(OUT)/dynamic.o: $(OUT)/mol.o
$(OUT)/sso.o: $(OUT)/mc.o
$(OUT)/molsim.o: $(OBJS)

$(OUT)/molsim: $(OUT)/molsim.o $(OBJS)
	$(REF_FC) $(FFLAGS) -o $@ $^

# ref_probe: our own test driver (oracle/ref_probe.F90) linked against the
# same objects.  molsim.o carries the reference's main program; its entry
# symbols are weakened in a copy so that the probe's own main wins.
PROBE := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))ref_probe.F90)

$(OUT)/molsim_nomain.o: $(OUT)/molsim.o
	objcopy --weaken-symbol=main --weaken-symbol=_QQmain $< $@

$(OUT)/ref_probe.o: $(PROBE) $(OBJS)
	cd $(OUT) && $(REF_FC) $(FFLAGS) -c $(PROBE) -o ref_probe.o

$(OUT)/ref_probe: $(OUT)/ref_probe.o $(OUT)/molsim_nomain.o $(OBJS)
	$(REF_FC) $(FFLAGS) -o $@ $^

clean:
	rm -rf $(OUT)

.PHONY: all clean
