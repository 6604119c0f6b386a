# Builds oracle/_ref/gpe_ref_driver from the reference's own sources where they
# lie (loader, parameters, units, context, initializer; the vendored
# nlohmann/json header) plus oracle/ref_driver.cpp.  TEST INFRASTRUCTURE: used
# only by tests/ to check this project's restatements.  The reference's solver
# sources are not built (they need fftw3.h, absent here).
REF ?= /root/reference
CXX ?= g++
OUT := _ref
SRCS := $(REF)/src/filemanager/file_manager.cpp $(REF)/src/parameters/parameters.cpp \
        $(REF)/src/utils/units/units.cpp $(REF)/src/context/context.cpp \
        $(REF)/src/initializer/initializer.cpp ref_driver.cpp
INC := -I$(REF)/src -I$(REF)/src/utils/output -I$(REF)/src/utils/units -I$(REF)/src/utils/types \
       -I$(REF)/lib

ifeq ($(wildcard $(REF)/src),)
# the reference sources are not on this machine: nothing to build
all:
	@true
else
all: $(OUT)/gpe_ref_driver
endif

$(OUT)/gpe_ref_driver: $(SRCS)
	mkdir -p $(OUT)
	$(CXX) -std=c++17 -O1 -w $(INC) $(SRCS) -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
