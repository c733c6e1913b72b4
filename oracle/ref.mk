# Builds the reference's own CPU emission path from its source files, in place
# under /root/reference (read-only), into oracle/_ref/.  Test infrastructure
# only: the result is a checker and a CPU baseline, never part of the product.
#
# The reference's CUDA translation units cannot be compiled here (no nvcc).
# The CPU sources below still name a few GPU-side symbols (virtual
# calc*GPU members, device-memory helpers); they are only reached when
# GlobalConfig::use_gpu is true, which the driver never sets.  They are left
# unresolved at link time (--unresolved-symbols=ignore-all) rather than
# replaced by stand-ins.
#
#   make -f oracle/ref.mk            (from the repository root)

REF      ?= /root/reference/src
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
CXX      ?= g++
CXXFLAGS := -std=c++11 -O2 -fopenmp -I$(REF)

SRCS := config/global_config.cpp \
        spectral_grid/spectral_grid.cpp spectral_grid/spectral_band.cpp \
        spectral_grid/spectral_band_convolve.cpp spectral_grid/spectral_band_integrate.cpp \
        spectral_grid/spectral_grid_interpolate.cpp spectral_grid/convert.cpp \
        observations/observations.cpp observations/observations_load_data.cpp \
        observations/observations_filter_response.cpp \
        additional/aux_functions.cpp additional/piecewise_poly.cpp \
        transport_coeff/opacity_species.cpp transport_coeff/sampled_data.cpp \
        transport_coeff/cross_section_file.cpp transport_coeff/transport_coeff.cpp \
        transport_coeff/species_rayleigh_cross_sections.cpp transport_coeff/h_m.cpp \
        forward_model/atmosphere/atmosphere.cpp \
        temperature/piecewise_poly_temperature.cpp \
        chemistry/isoprofile_chemistry.cpp chemistry/background_chemistry.cpp \
        chemistry/fixed_chemistry.cpp \
        radiative_transfer/short_characteristics.cpp \
        cloud_model/grey_cloud_model.cpp cloud_model/power_law_cloud_model.cpp \
        cloud_model/kh_cloud_model.cpp cloud_model/fixed_cloud_model.cpp \
        forward_model/stellar_spectrum/star_blackbody.cpp \
        forward_model/stellar_spectrum/star_file_spectrum.cpp

OBJS := $(addprefix $(OBJ)/,$(subst /,_,$(SRCS:.cpp=.o)))

all: $(OUT)/bear_ref_emission

$(OBJ):
	mkdir -p $@

define compile_rule
$(OBJ)/$(subst /,_,$(1:.cpp=.o)): $(REF)/$(1) | $(OBJ)
	$(CXX) $(CXXFLAGS) -c $$< -o $$@
endef
$(foreach s,$(SRCS),$(eval $(call compile_rule,$(s))))

# Every symbol left unresolved must be a GPU-side one (a *GPU member or a
# device-memory helper); the link fails otherwise.
$(OUT)/bear_ref_emission: oracle/ref_driver.cpp $(OBJS)
	@$(CXX) $(CXXFLAGS) oracle/ref_driver.cpp $(OBJS) -o $@.check 2>&1 \
	  | grep -o "undefined reference to \`[^']*'" | sort -u > $(OUT)/unresolved.txt || true
	@rm -f $@.check
	@if grep -v -E "GPU\(|Device<|Device\(" $(OUT)/unresolved.txt; then \
	  echo "non-GPU symbols unresolved (listed above)"; exit 1; fi
	$(CXX) $(CXXFLAGS) oracle/ref_driver.cpp $(OBJS) -o $@ -Wl,--unresolved-symbols=ignore-all

clean:
	rm -rf $(OUT)

.PHONY: all clean
