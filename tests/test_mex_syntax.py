"""The MEX gateway (3d-bi-anisotropic-phc_amd/mex/fame_mex.c) compiles against the C-ABI header.

MATLAB is not installed, so the gateway is checked with gcc -fsyntax-only against a
declarations-only stand-in for MATLAB's mex.h API; an entry point renamed or re-typed in
include/fame_abi.h without the gateway following fails here (implicit declarations are errors)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM_INC = "/opt/rocm/include"

MEX_API = """\
#include <stddef.h>
typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef struct { double real, imag; } mxComplexDouble;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
typedef enum { mxDOUBLE_CLASS, mxUINT8_CLASS, mxUINT64_CLASS } mxClassID;
void mexErrMsgIdAndTxt(const char*, const char*, ...);
int mxIsUint64(const mxArray*); int mxIsComplex(const mxArray*);
size_t mxGetNumberOfElements(const mxArray*); void* mxGetData(const mxArray*);
mxComplexDouble* mxGetComplexDoubles(const mxArray*); double* mxGetDoubles(const mxArray*);
mxArray* mxCreateDoubleMatrix(mwSize, mwSize, mxComplexity);
mxArray* mxCreateNumericMatrix(mwSize, mwSize, mxClassID, mxComplexity);
mxArray* mxCreateNumericArray(mwSize, const mwSize*, mxClassID, mxComplexity);
int mxGetString(const mxArray*, char*, mwSize); size_t mxGetM(const mxArray*); size_t mxGetN(const mxArray*);
double mxGetScalar(const mxArray*); void* mxMalloc(size_t); void* mxCalloc(size_t, size_t); void mxFree(void*);
void mexMakeMemoryPersistent(void*);
"""


@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.isdir(ROCM_INC), reason="needs gcc and ROCm headers")
def test_mex_gateway_compiles_against_header(tmp_path):
    (tmp_path / "mex.h").write_text(MEX_API)
    cmd = ["gcc", "-fsyntax-only", "-std=c99", "-Wall", "-Werror=implicit-function-declaration",
           "-Werror=incompatible-pointer-types", "-D__HIP_PLATFORM_AMD__", "-I", str(tmp_path),
           "-I", os.path.join(ROOT, "include"), "-I", ROCM_INC,
           os.path.join(ROOT, "3d-bi-anisotropic-phc_amd", "mex", "fame_mex.c")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
