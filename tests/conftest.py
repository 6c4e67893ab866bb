import os
import sys

import pytest

ROOT = os.path.dirname( os.path.dirname( os.path.abspath( __file__ ) ) )
for p in ( os.path.join( ROOT, "sdot-ffi_amd" ), os.path.join( ROOT, "oracle" ), ROOT ):
    if p not in sys.path:
        sys.path.insert( 0, p )


def pytest_configure( config ):
    config.addinivalue_line( "markers", "gpu: needs an MI355X ( runs the HIP library through its C ABI )" )


@pytest.fixture( scope = "session" )
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture( scope = "session" )
def native():
    """the HIP library through its C ABI; a GPU test fails ( it does not skip ) when it is missing"""
    from sdot import _native
    _native.lib()
    _native.device()
    return _native
