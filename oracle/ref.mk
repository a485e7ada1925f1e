# ORACLE — test infrastructure only. Compiles oracle/_ref from the reference's
# own sources, read in place (never copied). Invoked by oracle/Makefile only
# when the reference tree exists; outputs go to oracle/_ref/ only.
CXX ?= g++
REFERENCE ?= /root/reference
REF_SRC := $(REFERENCE)/src
REF_FLAGS := -O2 -std=c++20 -ffp-contract=off -fPIC -w -I$(REF_SRC)
OUT := _ref

all: $(OUT)/libref_kernels.so $(OUT)/libref_mesh.so $(OUT)/libref_devmesh.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/libref_kernels.so: ref_kernels.cpp | $(OUT)
	$(CXX) $(REF_FLAGS) -shared -o $@ ref_kernels.cpp

# Mesh construction, geometry and RCM renumbering need no generated code.
REF_MESH_SRCS := $(REF_SRC)/kutta/mesh/mesh.cpp $(REF_SRC)/kutta/mesh/topology.cpp \
                 $(REF_SRC)/kutta/io/generators.cpp
$(OUT)/libref_mesh.so: ref_mesh.cpp $(REF_MESH_SRCS) | $(OUT)
	$(CXX) $(REF_FLAGS) -shared -o $@ ref_mesh.cpp $(REF_MESH_SRCS)

# upload_mesh (fv/device_mesh.cpp) needs only the mesh code and the Device
# interface, which ref_devmesh.cpp implements over host memory.
REF_DEVMESH_SRCS := $(REF_SRC)/kutta/fv/device_mesh.cpp $(REF_MESH_SRCS)
$(OUT)/libref_devmesh.so: ref_devmesh.cpp $(REF_DEVMESH_SRCS) | $(OUT)
	$(CXX) $(REF_FLAGS) -shared -o $@ ref_devmesh.cpp $(REF_DEVMESH_SRCS) -lpthread

.PHONY: all
