# Compiles the reference's own drivers and test programs, unchanged, against the drop-in
# headers in include/ and links them with libqbf.so (no FFTW, no reference objects).
# Sources are read where they lie under /root/reference; outputs go to oracle/_ref/dropin/.
# Usage: make -f oracle/dropin.mk   (from the repository root)
REF ?= /root/reference
OUT = oracle/_ref/dropin
LIBDIR = $(abspath inhom-quantum-boltzmann-fourier-c_amd/lib)
CFLAGS = -O2 -std=gnu99 -w -Iinclude
LDFLAGS = -L$(LIBDIR) -Wl,-rpath,$(LIBDIR) -lqbf -lm
PROGS = main_hom main_inhom Cc_test Cd_test coll_step_test I1_test I2_test I2_test2 I3_test I4_test \
        lgwt_test sim_step_test slope_lim_test slope_lim_test2

all: $(addprefix $(OUT)/,$(PROGS))

$(OUT)/main_%: $(REF)/src/main_%.c $(LIBDIR)/libqbf.so
	mkdir -p $(OUT)
	gcc $(CFLAGS) -o $@ $< $(LDFLAGS)

$(OUT)/%: $(REF)/test/%.c $(LIBDIR)/libqbf.so
	mkdir -p $(OUT)
	gcc $(CFLAGS) -o $@ $< $(LDFLAGS)
