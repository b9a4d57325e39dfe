# ref.mk -- builds the reference's own C++ front-end (network construction, SpikingModel,
# monitors) and its Dummy backend from the sources where they lie under $(REF), with no
# change to any of them.  Outputs go to oracle/_ref/ only:
#   libspike_front.so    front-end + Dummy backend (every Spike/*.cpp except Context.cpp)
#   context_dummy.o      Spike/Backend/Context.cpp, default backend "Dummy"
#   context_hip.o        Spike/Backend/Context.cpp, default backend "HIP" (the reference's
#                        CMakeLists.txt:55 selects its CUDA backend the same way)
#   gen_connectivity     tests/golden/gen_connectivity.cpp on the Dummy front-end
#   Brunel10K_hip        Examples/Brunel10K.cpp unchanged, on the HIP backend
#   SimpleExample_hip    Examples/SimpleExample.cpp unchanged, on the HIP backend
#   frontend_net_hip     tests/frontend/frontend_net.cpp on the HIP backend
# The CUDA backend cannot be built here (no nvcc); Examples/VogelsAbbottNet.cpp includes
# Spike/Backend/CUDA/Helpers/ErrorCheck.hpp, which needs cuda.h, so it is not built.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
REF ?= /root/reference
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(ROOT)/oracle/_ref
OBJ := $(OUT)/obj
CXX ?= g++
CXXFLAGS := -std=c++11 -O2 -fPIC -DVERBOSE=1 -I$(REF) -w

FRONT_DIRS := Helpers Models Neurons Plasticity ActivityMonitor Synapses
FRONT_SRCS := $(foreach d,$(FRONT_DIRS),$(wildcard $(REF)/Spike/$(d)/*.cpp)) \
              $(shell find $(REF)/Spike/Backend/Dummy -name '*.cpp' | sort)
FRONT_OBJS := $(patsubst $(REF)/%.cpp,$(OBJ)/%.o,$(FRONT_SRCS))

HIP_LIB := $(ROOT)/spike_amd/_lib
BACKEND_LIB := $(HIP_LIB)/libspike_hip_backend.so
BACKEND_SRC := $(ROOT)/spike_amd/backend/HIPBackend.cpp

all: $(OUT)/libspike_front.so $(OUT)/context_dummy.o $(OUT)/context_hip.o $(OUT)/gen_connectivity

hip: all $(BACKEND_LIB) $(OUT)/Brunel10K_hip $(OUT)/SimpleExample_hip $(OUT)/frontend_net_hip

# The HIP backend (spike_amd/backend, this project's code) compiled against the Spike
# headers; it depends on the front-end library so that the factory maps it registers into
# are constructed before its registrars run.
$(BACKEND_LIB): $(BACKEND_SRC) $(ROOT)/spike_amd/backend/HIPBackend.hpp $(ROOT)/include/spike_hip.h $(OUT)/libspike_front.so $(HIP_LIB)/libspike_hip.so
	$(CXX) -std=c++14 -O2 -fPIC -shared -DVERBOSE=1 -I$(REF) -I$(ROOT)/include $(BACKEND_SRC) \
	  -L$(OUT) -lspike_front -L$(HIP_LIB) -lspike_hip \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../oracle/_ref' -o $@

$(OBJ)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libspike_front.so: $(FRONT_OBJS)
	$(CXX) -shared -o $@ $^

$(OUT)/context_dummy.o: $(REF)/Spike/Backend/Context.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/context_hip.o: $(REF)/Spike/Backend/Context.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) '-DSPIKE_DEFAULT_BACKEND="HIP"' -c $< -o $@

$(OUT)/gen_connectivity: $(ROOT)/tests/golden/gen_connectivity.cpp $(OUT)/libspike_front.so $(OUT)/context_dummy.o
	$(CXX) $(CXXFLAGS) -I$(REF)/Examples $< $(OUT)/context_dummy.o -L$(OUT) -lspike_front \
	  -Wl,-rpath,'$$ORIGIN' -o $@

$(OUT)/Brunel10K_hip: $(REF)/Examples/Brunel10K.cpp $(OUT)/libspike_front.so $(OUT)/context_hip.o $(BACKEND_LIB)
	$(CXX) $(CXXFLAGS) -I$(REF)/Examples $< $(OUT)/context_hip.o -L$(OUT) -lspike_front \
	  -L$(HIP_LIB) -Wl,--no-as-needed -lspike_hip_backend -Wl,--as-needed \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../spike_amd/_lib' -o $@

$(OUT)/SimpleExample_hip: $(REF)/Examples/SimpleExample.cpp $(OUT)/libspike_front.so $(OUT)/context_hip.o $(BACKEND_LIB)
	$(CXX) $(CXXFLAGS) -I$(REF)/Examples $< $(OUT)/context_hip.o -L$(OUT) -lspike_front \
	  -L$(HIP_LIB) -Wl,--no-as-needed -lspike_hip_backend -Wl,--as-needed \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../spike_amd/_lib' -o $@

$(OUT)/frontend_net_hip: $(ROOT)/tests/frontend/frontend_net.cpp $(OUT)/libspike_front.so $(OUT)/context_hip.o $(BACKEND_LIB)
	$(CXX) $(CXXFLAGS) -I$(REF)/Examples $< $(OUT)/context_hip.o -L$(OUT) -lspike_front \
	  -L$(HIP_LIB) -Wl,--no-as-needed -lspike_hip_backend -Wl,--as-needed \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../spike_amd/_lib' -o $@

.PHONY: all hip
