# Builds oracle/_ref/ref_golden from the reference's own compilable Fortran
# sources, compiled where they lie under /root/reference (nothing is copied).
# Run from oracle/_ref:  make -C oracle ref
# Outputs go to oracle/_ref/ only (git-ignored).
REF   ?= /root/reference/ufc_core
ifeq ($(origin FC),default)
FC    := flang
endif
FFLAGS ?= -O2 -cpp
HERE  := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT   := $(HERE)_ref

SRCS := $(REF)/L1_IF/Error/IF_Err_Def.f90 \
        $(REF)/L1_IF/Error/IF_Err_Brg.f90 \
        $(REF)/L1_IF/Precision/IF_Prec_Core.f90 \
        $(REF)/L2_NM/Matrix/NM_Mtx_Sparse.f90 \
        $(REF)/L2_NM/Solver/NM_Mtx_SpMVCSRTranspose.f90

OBJS := $(OUT)/IF_Err_Def.o $(OUT)/IF_Err_Brg.o $(OUT)/IF_Prec_Core.o $(OUT)/NM_Mtx_Sparse.o \
        $(OUT)/NM_Mtx_SpMVCSRTranspose.o

all: $(OUT)/ref_golden $(OUT)/ref_cantilever

$(OUT)/IF_Err_Def.o: $(REF)/L1_IF/Error/IF_Err_Def.f90
	$(FC) $(FFLAGS) -J $(OUT) -c $< -o $@
$(OUT)/IF_Err_Brg.o: $(REF)/L1_IF/Error/IF_Err_Brg.f90 $(OUT)/IF_Err_Def.o
	$(FC) $(FFLAGS) -J $(OUT) -c $< -o $@
$(OUT)/IF_Prec_Core.o: $(REF)/L1_IF/Precision/IF_Prec_Core.f90 $(OUT)/IF_Err_Brg.o
	$(FC) $(FFLAGS) -J $(OUT) -c $< -o $@
$(OUT)/NM_Mtx_Sparse.o: $(REF)/L2_NM/Matrix/NM_Mtx_Sparse.f90 $(OUT)/IF_Prec_Core.o
	$(FC) $(FFLAGS) -J $(OUT) -c $< -o $@

$(OUT)/NM_Mtx_SpMVCSRTranspose.o: $(REF)/L2_NM/Solver/NM_Mtx_SpMVCSRTranspose.f90 $(OUT)/IF_Prec_Core.o
	$(FC) $(FFLAGS) -J $(OUT) -c $< -o $@

$(OUT)/ref_golden: $(HERE)ref_golden.f90 $(OBJS)
	$(FC) $(FFLAGS) -J $(OUT) -I $(OUT) $(HERE)ref_golden.f90 $(OBJS) -o $@

# second-opinion CPU timing of the small cantilever (reference CSR assembly and SpMV)
$(OUT)/ref_cantilever: $(HERE)ref_cantilever.f90 $(OBJS)
	$(FC) $(FFLAGS) -J $(OUT) -I $(OUT) $(HERE)ref_cantilever.f90 $(OBJS) -o $@

.PHONY: all
