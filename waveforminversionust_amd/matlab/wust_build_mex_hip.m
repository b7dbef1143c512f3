function wust_build_mex_hip(rocm)
    % Build the wust_hip_mex gateway of the MI355X engine (counterpart of Runtime/matlab/wust_build_mex.m).
    % The engine library (waveforminversionust_amd/_lib/libwust_hip.so) must be built first, e.g. with
    % python -c "import __graft_entry__ as g; g.build()" at the repository root.
    if nargin < 1
        rocm = '/opt/rocm';
    end
    here = fileparts(mfilename('fullpath'));
    root = fileparts(fileparts(here));
    lib = fullfile(root, 'waveforminversionust_amd', '_lib');
    mex('-R2018a', '-D__HIP_PLATFORM_AMD__', ['-I' fullfile(root, 'include')], ['-I' fullfile(rocm, 'include')], ...
        ['-L' lib], '-lwust_hip', ['-L' fullfile(rocm, 'lib')], '-lamdhip64', ...
        ['LDFLAGS=$LDFLAGS -Wl,-rpath,' lib ' -Wl,-rpath,' fullfile(rocm, 'lib')], ...
        '-outdir', here, fullfile(here, 'wust_hip_mex.cpp'));
end
