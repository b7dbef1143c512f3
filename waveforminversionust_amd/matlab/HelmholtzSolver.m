classdef HelmholtzSolver < handle
    % Drop-in for Runtime/solver/HelmholtzSolver.m (gpu backend) on the MI355X engine.
    % Same constructor HelmholtzSolver(x, y, vel, atten, f, signConvention, a0, L_PML, options) and
    % the same checks and WUST:* identifiers (HelmholtzSolver.m:32-52); solve(src, adjoint) and
    % massApply(v, adjoint) keep their meaning (:174-220).  The operator and its factors live on
    % the GPU behind a uint64 handle of the wust_hip_mex gateway (property engine) (MATLAB gpuArrays cannot hold HIP
    % memory); the handle is freed when the object is deleted.  Fields cross to and from the host
    % as complex double arrays.  options.mode ('refined' default, 'c64', 'plain'), options.factor
    % ('auto' or 'gauss-jordan'), options.tol (1e-10) and options.max_iter (30) select the arithmetic.

    properties (SetAccess = private)
        backend
        engine = uint64(0)   % wust_hip_mex handle of the device operator
        ny
        nx
    end

    properties
        assembly_seconds
        factor_seconds
        massD
        massE
        tol = 1e-10
        max_iter = 30
    end

    methods
        function obj = HelmholtzSolver(x, y, vel, atten, f, signConvention, a0, L_PML, options)
            assert(nargin == 9 && isstruct(options) && ...
                   all(isfield(options, {'backend', 'stencil_bounds'})), ...
                   'WUST:InvalidInput', 'Explicit backend and fixed stencil_bounds are required');
            bounds = options.stencil_bounds;
            assert(isnumeric(bounds) && isreal(bounds) && numel(bounds) == 2 && ...
                   all(isfinite(bounds)) && bounds(1) > 0 && bounds(2) > bounds(1), ...
                   'WUST:InvalidInput', 'stencil_bounds must be a positive ordered [min,max] pair');
            assert(ismember(string(options.backend), ["cpu", "gpu"]), ...
                   'WUST:InvalidInput', 'backend must be cpu or gpu');
            assert(strcmp(string(options.backend), "gpu"), ...
                   'WUST:InvalidInput', 'this engine implements the gpu backend only');
            assert(isequal(size(vel), [numel(y), numel(x)]) && min(size(vel)) >= 3 && ...
                   isequal(size(atten), size(vel)) && all(isfinite(vel(:)) & vel(:) > 0), ...
                   'WUST:InvalidInput', 'Expected a positive finite [ny,nx] medium');
            assert(all(diff(x) > 0) && all(diff(y) > 0) && ...
                   max(abs(diff(x) - mean(diff(x)))) < mean(diff(x)) * 1e-7 && ...
                   max(abs(diff(y) - mean(diff(y)))) < mean(diff(y)) * 1e-7, ...
                   'WUST:InvalidInput', 'Uniform increasing grids required');
            mode = 'refined';
            if isfield(options, 'mode') && ~isempty(options.mode)
                mode = char(options.mode);
            end
            if isfield(options, 'factor') && strcmp(char(options.factor), 'gauss-jordan')
                mode = [mode '-gauss-jordan'];
            end
            if isfield(options, 'tol') && ~isempty(options.tol)
                obj.tol = double(options.tol);
            end
            if isfield(options, 'max_iter') && ~isempty(options.max_iter)
                obj.max_iter = double(options.max_iter);
            end
            dispersion = zeros(0, 0);
            if isfield(options, 'dispersion') && ~isempty(options.dispersion)
                dispersion = [double(options.dispersion.time_step_s), ...
                              double(options.dispersion.reference_speed_mps)];
            end
            obj.backend = 'gpu';
            [obj.ny, obj.nx] = size(vel);
            [obj.engine, info] = wust_hip_mex('create', double(x(:)).', double(y(:)).', double(vel), ...
                                              double(atten), double(f), double(signConvention), ...
                                              double(a0), double(L_PML), double(bounds(:)).', mode, dispersion);
            obj.assembly_seconds = info(1) / 1e3;
            obj.factor_seconds = info(2) / 1e3;
            obj.massD = info(4);
            obj.massE = info(5);
        end

        function delete(obj)
            if obj.engine ~= 0
                wust_hip_mex('destroy', obj.engine);
                obj.engine = uint64(0);
            end
        end

        function [wvfield, virtSrcs] = solve(obj, src, adjoint)
            % Solve H u = src or H' u = src; virtSrcs = (C .* q') .* u (HelmholtzSolver.m:174-203).
            src = complex(double(gather(src)));
            if nargout > 1
                [wvfield, virtSrcs] = wust_hip_mex('solve', obj.engine, src, double(adjoint), obj.tol, obj.max_iter);
            else
                wvfield = wust_hip_mex('solve', obj.engine, src, double(adjoint), obj.tol, obj.max_iter);
            end
        end

        function out = massApply(obj, v, adjoint)
            % Exact derivative of the optimized nine-point mass stencil (HelmholtzSolver.m:205-220).
            out = wust_hip_mex('mass', obj.engine, complex(double(gather(v))), double(adjoint));
        end
    end
end
