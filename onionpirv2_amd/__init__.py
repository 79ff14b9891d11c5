"""MI355X-native OnionPIRv2 server answer path.

  onionpirv2_amd.gpu   ctypes binding of the C ABI (include/onionpir_gpu.h)
  onionpirv2_amd.host  the reference's PirParams / PirClient / PirServer API
                       (C++ host library; the server runs on the GPU)
"""
