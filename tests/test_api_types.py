"""The C++ API's header-only parts, through the compiled driver
tests/cpp/api_types.cpp (built next to the library by the csrc Makefile):
  * num_loc / indxg2p / indxg2l / indxl2g against a brute-force walk of the
    block-cyclic distribution (include/ddla/ddla_desc.h:10-30 of the
    reference), for block sizes 1-64, 1-5 processes, every source process;
  * ddla::vector<T> (include/ddla/ddla_vector.h:27-110 of the reference) on
    the GPU handle: growth keeps the values, shrinking / clear() keep the
    allocation, moves transfer it, reset() releases it, bad arguments throw;
  * the handle helpers (include/ddla/ddla_handle_t.h:146-334 of the
    reference): pinned host memory, ddlaMemsetTri / ddlaMemcpyTri in every
    direction against element-wise expectations (both triangles, with and
    without the diagonal, padding and the other triangle untouched, a 3-byte
    element), events, device queries, the rank <-> grid maps and the argument
    rules.
"""
import os
import subprocess

import pytest

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "api_types")


def run_driver(*args):
    assert os.path.exists(DRIVER), f"{DRIVER} is not built (make -C libddla_amd/csrc)"
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failure(s)" in out.stdout
    return out.stdout


def test_block_cyclic_index_maps():
    run_driver("cpu")


@pytest.mark.gpu
def test_vector_and_handle_helpers_on_gpu():
    print("case ddla::vector and the handle helpers on the GPU handle", flush=True)
    run_driver()


MPI_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "api_mpi")


def run_mpi_driver(*args):
    assert os.path.exists(MPI_DRIVER), f"{MPI_DRIVER} is not built (make -C libddla_amd/csrc)"
    out = subprocess.run([MPI_DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failure(s)" in out.stdout


def test_communicator_overloads_of_ddlaset():
    """ddlaSet(handle, const MPI_Comm&, ...) (include/ddla/ddla_handle_t.h:58-64
    of the reference) compiled as an MPI caller compiles: more than one rank is
    refused before any device work, one rank forwards to the 1 x 1 grid, and
    integer grid shapes still pick the grid-shape overload."""
    run_mpi_driver()


@pytest.mark.gpu
def test_communicator_overloads_set_up_gpu_handles():
    print("case ddlaSet through a one-rank communicator on the GPU handle", flush=True)
    run_mpi_driver("gpu")
