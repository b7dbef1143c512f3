function [loss, gradient, cache] = wust_oracle(c, obs, fi, cfg)
    % Drop-in for Runtime/matlab/wust_oracle.m on the MI355X engine: the loss, slowness gradient and
    % (nargout 3) retained curvature state of one frequency in a single engine call (forward solves,
    % FP64 source fit a = (p' d) / (p' p), adjoint solves reusing the factors, mass-stencil gradient).
    % Same arguments, nargout contract and WUST:* identifiers; cache.solver is the engine's
    % HelmholtzSolver, whose handle keeps the forward fields and scales for directional_curvature.
    % cfg.mode / cfg.tol / cfg.max_iter / cfg.factor select the engine arithmetic as in HelmholtzSolver.
    need_gradient = nargout > 1;
    retain_state = nargout > 2;
    ny = obs.shape_yx(1);
    nx = obs.shape_yx(2);
    options = struct('backend', cfg.backend, 'stencil_bounds', cfg.stencil_bounds, 'dispersion', []);
    if strcmp(cfg.wavenumber, 'kwave-ldr9')
        options.dispersion = cfg.dispersion;
    end
    for name = {'mode', 'tol', 'max_iter', 'factor'}
        if isfield(cfg, name{1})
            options.(name{1}) = cfg.(name{1});
        end
    end
    solver = HelmholtzSolver(obs.x_m, obs.y_m, double(c), zeros(ny, nx), obs.frequencies_hz(fi), -1, ...
                             cfg.pml_strength, cfg.pml_m, options);
    timer = tic;
    [loss, gradient, scales] = wust_hip_mex('oracle', solver.engine, double(obs.tx_index(:)).', ...
                                            double(obs.rx_index(:)).', complex(double(obs.Y(:, :, fi))), ...
                                            logical(obs.mask(:, :, fi)), double(cfg.source_batch_size), ...
                                            double(need_gradient), double(retain_state), solver.tol, ...
                                            solver.max_iter);
    seconds = toc(timer);
    assert(isfinite(loss) && (~need_gradient || all(isfinite(gradient(:)))), ...
           'WUST:NumericalFailure', 'Nonfinite objective/gradient');
    if retain_state
        cache = struct('solver', solver, 'virtual', {{}}, 'scales', {{scales}}, ...
                       'forward_seconds', seconds, 'adjoint_seconds', 0, ...
                       'forward_rhs', obs.n_tx, 'adjoint_rhs', obs.n_tx);
    end
end
