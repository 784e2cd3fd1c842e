{
  "targets": [
    {
      "target_name": "clay_hip",
      "sources": ["clay_napi.cc"],
      "include_dirs": ["../include"],
      "cflags_cc": ["-std=c++17", "-O2"],
      "libraries": ["-L<(module_root_dir)/../clay-engine_amd", "-lclay_hip",
                    "-Wl,-rpath,<(module_root_dir)/../clay-engine_amd"]
    }
  ]
}
